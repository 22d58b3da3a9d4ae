"""ctypes prototypes of the exported C-ABI (include/yael_mi355x.h)."""
import ctypes as C

f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int)
u16p = C.POINTER(C.c_uint16)
vp = C.c_void_p
L = C.c_long
I = C.c_int

# name -> (restype, argtypes)
PROTOS = {
    # yael C API (nn.h / kmeans.h)
    "knn_full": (None, [I, I, I, I, I, vp, vp, vp, vp, vp]),
    "knn_full_thread": (None, [I, I, I, I, I, vp, vp, vp, vp, vp, I]),
    "nn": (C.c_double, [I, I, I, vp, vp, vp]),
    "nn_thread": (C.c_double, [I, I, I, vp, vp, vp, I]),
    "knn": (vp, [I, I, I, I, vp, vp, vp]),
    "knn_thread": (vp, [I, I, I, I, vp, vp, vp, I]),
    "yael_knn_precision": (I, [I]),
    # hkm.h
    "hkm_learn": (vp, [I, I, I, I, vp, I, I, I, vp]),
    "hkm_delete": (None, [vp]),
    "hkm_quantize": (None, [vp, I, vp, vp]),
    "hkm_write": (None, [C.c_char_p, vp]),
    "hkm_read": (vp, [C.c_char_p]),
    "hkm_get_centroids": (vp, [vp, I, I]),
    "yael_hkm_learn_checked": (I, [I, I, I, I, vp, I, I, I, vp, vp]),
    "yael_hkm_quantize_checked": (I, [vp, I, vp, vp]),
    "compute_cross_distances": (None, [I, I, I, vp, vp, vp]),
    "compute_cross_distances_nonpacked": (None, [I, I, I, vp, I, vp, I, vp, I]),
    "compute_cross_distances_thread": (None, [I, I, I, vp, vp, vp, I]),
    "compute_cross_distances_alt": (None, [I, I, I, I, vp, vp, vp]),
    "compute_cross_distances_alt_nonpacked": (None, [I, I, I, I, vp, I, vp, I, vp, I]),
    "compute_cross_distances_alt_thread": (None, [I, I, I, I, vp, vp, vp, I]),
    "kmeans": (C.c_float, [I, I, I, I, vp, I, L, I, vp, vp, vp, vp]),
    "clustering_kmeans": (vp, [I, I, vp, I, I, C.c_double]),
    "clustering_kmeans_assign": (vp, [I, I, vp, I, I, C.c_double, C.POINTER(vp)]),
    "clustering_kmeans_assign_with_score": (vp, [I, I, vp, I, I, C.c_double, I, f64p,
                                                 C.POINTER(vp)]),
    # status-returning variants (same arguments, + float* result for kmeans)
    "yael_knn_full_checked": (I, [I, I, I, I, I, vp, vp, vp, vp, vp]),
    "yael_cross_distances_checked": (I, [I, I, I, I, vp, I, vp, I, vp, I]),
    "yael_kmeans_checked": (I, [I, I, I, I, vp, I, L, I, vp, vp, vp, vp, f32p]),
    "yael_knn_reorder_shortlist_checked": (I, [I, I, I, I, vp, vp, vp, vp]),
    "yael_kmeans_device": (I, [I, I, I, I, vp, I, L, I, vp, vp, vp, vp, f32p]),
    "yael_kmeans_batched": (I, [I, I, vp, I, I, vp, I, vp, vp, vp, vp, vp, vp]),
    "yael_kmeans_batched_device": (I, [I, I, vp, I, I, vp, I, vp, vp, vp, vp, vp, vp, vp, I,
                                       vp]),
    "yael_kmeans_last_stats": (I, [C.POINTER(L), C.POINTER(L), C.POINTER(L)]),
    "yael_kmeans_subspaces_device": (I, [I, I, I, I, I, vp, I, vp, vp, vp, vp, vp, vp, vp,
                                         vp]),
    "yk_copy_cols": (I, [vp, I, vp, I, L, I]),
    "yk_copy_kernel": (I, [vp, vp, C.c_size_t]),
    "yk_bf16_bytes": (L, [L, I, I, I]),
    "yk_synth": (I, [vp, L, I, I, I, C.c_float, C.c_float, vp, I, C.c_float, I, C.c_ulonglong,
                     I]),
    "yk_gather_rows": (I, [vp, vp, L, I, vp]),
    "yk_rerank_shortlist": (I, [L, I, I, vp, vp, vp, vp]),
    "knn_reorder_shortlist": (None, [I, I, I, I, vp, vp, vp, vp]),
    "knn_recompute_exact_dists": (None, [I, I, I, I, vp, vp, I, vp, vp, vp]),
    "yael_knn_recompute_exact_dists_checked": (I, [I, I, I, I, vp, vp, I, vp, vp, vp]),
    "yk_recompute_exact_dists": (I, [L, L, I, I, vp, vp, I, vp, vp, vp, vp]),
    # device-level C-ABI (yk.h)
    "yk_last_error": (C.c_char_p, []),
    "yk_set_error_c": (I, [I, C.c_char_p]),
    "yk_init": (I, []),
    "yk_ld": (I, [I]),
    "yk_malloc": (I, [C.POINTER(vp), C.c_size_t]),
    "yk_free": (I, [vp]),
    "yk_upload_vectors": (I, [vp, vp, L, I]),
    "yk_download_vectors": (I, [vp, vp, L, I]),
    "yk_memcpy_h2d": (I, [vp, vp, C.c_size_t]),
    "yk_host_alloc": (I, [C.POINTER(vp), C.c_size_t]),
    "yk_host_free": (I, [vp]),
    "yk_kpp_update": (I, [vp, L, I, vp, L, vp, I, vp]),
    "yk_hkm_quantize": (I, [vp, vp, L, I, I, I, vp, vp, vp]),
    "yk_memcpy_d2h": (I, [vp, vp, C.c_size_t]),
    "yk_memcpy_d2d": (I, [vp, vp, C.c_size_t]),
    "yk_memset": (I, [vp, I, C.c_size_t]),
    "yk_sync": (I, []),
    "yk_norms": (I, [vp, L, I, vp, vp]),
    "yk_knn_f32": (I, [I, L, L, I, I, vp, vp, vp, vp, vp, vp, vp]),
    "yk_knn_f32_noalloc": (I, [I, L, L, I, I, vp, vp, vp, vp, vp, vp, vp]),
    "yk_cross_distances": (I, [I, L, L, I, vp, vp, vp, vp, vp, L]),
    "yk_bf16_ld": (I, [I, I]),
    "yk_to_bf16": (I, [vp, L, I, I, I, vp, vp, vp, vp]),
    "yk_knn_bf16": (I, [I, L, L, I, I, I, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32p]),
    "yk_pair_dist_f64": (I, [I, L, I, I, vp, vp, vp, vp]),
    "yk_kmeans_workspace_size": (C.c_size_t, [L, I, I]),
    "yk_kmeans_update": (I, [L, I, I, vp, vp, vp, vp, I, vp]),
    "yk_sum_f64": (I, [vp, L, f64p, vp]),
    "yk_timer_start": (I, []),
    "yk_timer_stop": (I, [f32p]),
    "yk_prof_enable": (I, [I]),
    "yk_prof_get": (I, [I, f64p, C.POINTER(L)]),
    "yk_prof_reset": (I, []),
    "yk_tune_get": (I, [C.c_char_p, i32p]),
    "yk_tune_put": (None, [C.c_char_p, I]),
    "yk_tune_enabled": (I, []),
    "yk_device_info": (I, [C.c_char_p, I]),
    "yk_knn_plan": (I, [L, L, I, I, i32p, i32p, i32p]),
}


def declare(lib):
    for name, (res, args) in PROTOS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
