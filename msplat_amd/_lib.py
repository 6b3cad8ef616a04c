"""ctypes binding of libmsplat_core.so (the C ABIs in include/*.h).

The library is built in-tree by msplat_amd.build; importing this module without it fails
loudly: there is no CPU fallback on the product path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmsplat_core.so")


class MsplatCamera(C.Structure):
    _fields_ = [
        ("viewmat", C.c_float * 16),
        ("projmat", C.c_float * 16),
        ("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float),
        ("cam_pos", C.c_float * 3),
        ("width", C.c_int32), ("height", C.c_int32),
    ]


class MsplatStepParams(C.Structure):
    _fields_ = [
        ("step_size", C.c_float * 6),
        ("bc2_sqrt", C.c_float * 6),
        ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
        ("ssim_weight", C.c_float),
        ("loss_inv_n", C.c_float),
        ("op_reg", C.c_float), ("sc_reg", C.c_float),
        ("reg_invis_scale", C.c_float),
        ("background", C.c_float * 3),
        ("noise_lr", C.c_float), ("xyz_lr", C.c_float), ("max_disp", C.c_float),
        ("sgld_seed", C.c_uint32),
        ("sgld_enable", C.c_int32),
        ("degree", C.c_int32),
        ("degrees_to_use", C.c_int32),
    ]


class MsplatDeviceState(C.Structure):
    _fields_ = [
        ("N", C.c_int32), ("degree", C.c_int32),
        ("means", C.c_void_p), ("scales", C.c_void_p), ("quats", C.c_void_p),
        ("f_dc", C.c_void_p), ("f_rest", C.c_void_p), ("opac", C.c_void_p),
        ("m", C.c_void_p * 6), ("v", C.c_void_p * 6),
        ("sh_bf16", C.c_int32), ("f_dc_bf16", C.c_void_p), ("f_rest_bf16", C.c_void_p),
    ]


class MsplatConfig(C.Structure):
    """Layout of the reference's MsplatConfig (core/include/msplat_c_api.h:16-30)."""
    _fields_ = [
        ("iterations", C.c_int), ("shDegree", C.c_int), ("shDegreeInterval", C.c_int),
        ("ssimWeight", C.c_float), ("capMax", C.c_int), ("noiseLr", C.c_float),
        ("opacityReg", C.c_float), ("scaleReg", C.c_float), ("cullRadius", C.c_float),
        ("keepCrs", C.c_bool), ("downscaleFactor", C.c_float), ("bgColor", C.c_float * 3),
    ]


class MsplatStats(C.Structure):
    _fields_ = [("iteration", C.c_int), ("splatCount", C.c_int), ("msPerStep", C.c_float)]


class MsplatEvalMetrics(C.Structure):
    _fields_ = [("psnr", C.c_float), ("ssim", C.c_float), ("l1", C.c_float), ("numTest", C.c_int),
                ("numGaussians", C.c_int)]


class MsplatPixelBuffer(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_float)), ("width", C.c_int), ("height", C.c_int)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP extension first (python -m msplat_amd.build)")
    _lib = bind(LIB_PATH)
    return _lib


def bind(path: str) -> C.CDLL:
    """Loads a library exporting the C ABI at `path` and declares every signature."""
    L = C.CDLL(path)
    vp, i32, i64, f32, f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_float, C.c_double
    fp = C.POINTER(C.c_float)
    sig = {
        # kernel-level ABI (msplat_kernels.h)
        "msplat_dev_count": ([], C.c_int),
        "msplat_dev_alloc": ([C.c_size_t], vp),
        "msplat_dev_free": ([vp], None),
        "msplat_dev_upload": ([vp, vp, C.c_size_t], C.c_int),
        "msplat_dev_download": ([vp, vp, C.c_size_t], C.c_int),
        "msplat_dev_memset": ([vp, C.c_int, C.c_size_t], C.c_int),
        "msplat_dev_copy_kernel": ([vp, vp, C.c_size_t], C.c_int),
        "msplat_set_list_bucket_waves": ([C.c_int], C.c_int),
        "msplat_list_bucket_waves_for": ([C.c_int, i64, C.c_int], C.c_int),
        "msplat_engine_create": ([], vp),
        "msplat_engine_destroy": ([vp], None),
        "msplat_engine_train_step": ([vp, C.POINTER(MsplatDeviceState), C.POINTER(MsplatCamera), vp,
                                      C.POINTER(MsplatStepParams), C.c_int], C.c_int),
        "msplat_engine_train_step_masked": ([vp, C.POINTER(MsplatDeviceState), C.POINTER(MsplatCamera), vp, vp,
                                             C.POINTER(MsplatStepParams), C.c_int], C.c_int),
        "msplat_engine_render": ([vp, C.POINTER(MsplatDeviceState), C.POINTER(MsplatCamera), C.c_int, fp], C.c_int),
        "msplat_render_noalloc": ([C.POINTER(MsplatDeviceState), C.POINTER(MsplatCamera), C.c_int, fp, vp, vp, vp,
                                   vp], C.c_int),
        "msplat_engine_buffer": ([vp, C.c_char_p, C.POINTER(vp), C.POINTER(C.c_size_t)], C.c_int),
        "msplat_engine_sync": ([vp], C.c_int),
        "msplat_engine_read": ([vp, C.c_char_p, vp, C.c_size_t], i64),
        "msplat_engine_stats": ([vp, C.POINTER(i64)], C.c_int),
        "msplat_engine_loss": ([vp], f64),
        "msplat_engine_stream": ([vp], vp),
        "msplat_engine_mark": ([vp], None),
        "msplat_engine_elapsed_ms": ([vp], f32),
        "msplat_engine_set_stage_timing": ([vp, C.c_int], None),
        "msplat_engine_stage_times": ([vp, C.POINTER(f64), C.c_int, C.POINTER(i64)], C.c_int),
        "msplat_engine_stage_name": ([C.c_int], C.c_char_p),
        "msplat_engine_reset_stage_times": ([vp], None),
        # scene helpers (msplat_api_ext.h)
        "msplat_camera_prepare": ([fp, f32, f32, f32, f32, C.c_int, C.c_int, C.POINTER(MsplatCamera)], None),
        "msplat_synth_gaussians": ([C.c_int, C.c_uint64, C.c_int, C.c_int, fp, fp, fp, fp, fp, fp], C.c_int),
        "msplat_synth_cameras": ([C.c_uint64, C.c_int, f32, f32, fp], C.c_int),
        "msplat_quantize_rgb": ([fp, i64, C.POINTER(C.c_uint8)], None),
        "msplat_camera_from_fov": ([fp, f32, C.c_int, C.c_int, C.POINTER(MsplatCamera)], None),
        # the msplat C API (msplat_c_api.h)
        "msplat_dataset_create": ([C.c_char_p, f32, C.c_bool, C.c_int], vp),
        "msplat_dataset_destroy": ([vp], None),
        "msplat_dataset_num_train": ([vp], C.c_int),
        "msplat_dataset_num_test": ([vp], C.c_int),
        "msplat_dataset_camera_pose": ([vp, C.c_int, fp], None),
        "msplat_trainer_create": ([vp, MsplatConfig], vp),
        "msplat_trainer_destroy": ([vp], None),
        "msplat_trainer_step": ([vp], MsplatStats),
        "msplat_trainer_train": ([vp], None),
        "msplat_trainer_evaluate": ([vp], MsplatEvalMetrics),
        "msplat_trainer_render": ([vp, C.c_int, C.c_bool], MsplatPixelBuffer),
        "msplat_trainer_render_pose": ([vp, fp, C.c_int], MsplatPixelBuffer),
        "msplat_trainer_render_pose_to_buffer": ([vp, fp, C.c_int, C.POINTER(C.c_uint8), C.POINTER(C.c_int),
                                                  C.POINTER(C.c_int)], None),
        "msplat_trainer_export_ply": ([vp, C.c_char_p], None),
        "msplat_trainer_export_splat": ([vp, C.c_char_p], None),
        "msplat_trainer_save_checkpoint": ([vp, C.c_char_p], None),
        "msplat_trainer_load_checkpoint": ([vp, C.c_char_p], C.c_int),
        "msplat_trainer_splat_count": ([vp], C.c_int),
        "msplat_trainer_iteration": ([vp], C.c_int),
        "msplat_set_metallib_path": ([C.c_char_p], None),
        "msplat_sync": ([], None),
        "msplat_cleanup": ([], None),
        "msplat_viewer_create": ([C.c_char_p, C.c_int, C.c_int, f32], vp),
        "msplat_viewer_resize": ([vp, C.c_int, C.c_int], None),
        "msplat_viewer_set_fov": ([vp, f32], None),
        "msplat_viewer_set_background": ([vp, f32, f32, f32], None),
        "msplat_viewer_render": ([vp, fp, C.POINTER(C.c_uint8)], None),
        "msplat_viewer_splat_count": ([vp], C.c_int),
        "msplat_viewer_get_size": ([vp, C.POINTER(C.c_int), C.POINTER(C.c_int)], None),
        "msplat_viewer_render_buffer": ([vp, fp, C.POINTER(C.c_int), C.POINTER(C.c_int)], vp),
        "msplat_viewer_bounds": ([vp, fp, fp], None),
        "msplat_viewer_density_center": ([vp, fp], None),
        "msplat_viewer_command_queue": ([vp], vp),
        "msplat_viewer_destroy": ([vp], None),
        # C API extensions (msplat_api_ext.h)
        "msplat_dataset_create_empty": ([], vp),
        "msplat_dataset_add_camera": ([vp, fp, f32, f32, f32, f32, C.c_int, C.c_int, C.POINTER(C.c_uint8),
                                       C.c_bool], C.c_int),
        "msplat_dataset_camera_info": ([vp, C.c_int, C.c_bool, fp, fp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                        C.POINTER(C.c_uint8)], C.c_int),
        "msplat_dataset_set_init_gaussians": ([vp, C.c_int, C.c_int, fp, fp, fp, fp, fp, fp], C.c_int),
        "msplat_dataset_set_mask": ([vp, C.c_int, C.c_bool, C.POINTER(C.c_uint8)], C.c_int64),
        "msplat_dataset_load_masks": ([vp, C.c_char_p], C.c_int64),
        # host-side ingestion (msplat_io.h)
        "msplat_image_decode": ([C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), vp], C.c_int),
        "msplat_image_resize_area": ([vp, C.c_int, C.c_int, vp, C.c_int, C.c_int], C.c_int),
        "msplat_image_undistort": ([vp, C.c_int, C.c_int, f32, f32, f32, f32, vp, vp, C.POINTER(C.c_int),
                                    C.POINTER(C.c_int), vp], C.c_int),
        "msplat_scene_load": ([C.c_char_p, C.c_char_p], vp),
        "msplat_scene_free": ([vp], None),
        "msplat_scene_num_cameras": ([vp], C.c_int),
        "msplat_scene_camera": ([vp, C.c_int, vp, vp, vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_char_p, C.c_int,
                                 C.c_char_p, C.c_int], C.c_int),
        "msplat_scene_transform": ([vp, C.POINTER(C.c_float), vp], C.c_int),
        "msplat_scene_num_points": ([vp], C.c_int64),
        "msplat_scene_points": ([vp, vp, vp], C.c_int),
        "msplat_scene_load_image": ([vp, C.c_int, f32], C.c_int),
        "msplat_scene_image": ([vp, C.c_int, vp], C.c_int),
        "msplat_scene_load_mask": ([vp, C.c_int, C.c_char_p, vp], C.c_int64),
        "msplat_knn_mean_distance": ([vp, C.c_int64, C.c_int, vp], C.c_int),
        "msplat_seed_gaussians": ([vp, vp, C.c_int64, C.c_int, f32, f32, vp, vp, vp, vp, vp, vp], C.c_int),
        "msplat_trainer_step_params": ([vp, C.c_int, C.c_int, C.c_int, C.POINTER(MsplatStepParams)], C.c_int),
        "msplat_trainer_last_camera": ([vp], C.c_int),
        "msplat_trainer_step_batch": ([vp, C.c_int], MsplatStats),
        "msplat_trainer_last_batch": ([vp, C.POINTER(i32), C.POINTER(f64), C.c_int], C.c_int),
        "msplat_trainer_batch_info": ([vp, C.POINTER(C.c_int64)], C.c_int),
        "msplat_trainer_proposal_info": ([vp, C.POINTER(f64)], C.c_int),
        "msplat_trainer_autotune_info": ([vp, C.POINTER(i32), fp, C.c_int, C.POINTER(i32)],
                                         C.c_int),
        "msplat_trainer_proposal_params": ([vp, fp, C.c_int], C.c_int),
        "msplat_trainer_proposal_read": ([vp, C.c_char_p, vp, C.c_int64], C.c_int64),
        "msplat_mlp_packed_bytes": ([], C.c_size_t),
        "msplat_mlp_pack": ([vp, vp], C.c_int),
        "msplat_mlp_forward": ([C.c_int, vp, vp, vp], C.c_int),
        "msplat_mlp_backward": ([C.c_int, vp, vp, vp, vp, vp], C.c_int),
        "msplat_trainer_last_loss": ([vp], f64),
        "msplat_trainer_read_group": ([vp, C.c_int, C.c_int, fp, i64], i64),
        "msplat_trainer_run_mcmc": ([vp, C.c_int, C.POINTER(i32)], C.c_int),
        "msplat_trainer_set_option": ([vp, C.c_char_p, f64], C.c_int),
        "msplat_trainer_optim_state": ([vp, C.POINTER(C.c_int), fp], C.c_int),
        "msplat_trainer_engine": ([vp], vp),
        "msplat_eq9_device": ([C.c_int, fp, C.POINTER(C.c_int), fp, fp], C.c_int),
    }
    for name, (argtypes, restype) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = argtypes
        fn.restype = restype
    return L


def default_config() -> MsplatConfig:
    """msplat_default_config() (a static inline in the header, restated for ctypes)."""
    c = MsplatConfig()
    c.iterations, c.shDegree, c.shDegreeInterval, c.ssimWeight = 30000, 3, 1000, 0.2
    c.capMax, c.noiseLr, c.opacityReg, c.scaleReg, c.cullRadius = 1000000, 5e5, 0.01, 0.01, 3.0
    c.keepCrs, c.downscaleFactor = False, 1.0
    c.bgColor[0], c.bgColor[1], c.bgColor[2] = 0.6130, 0.0101, 0.3984
    return c


def fptr(a):
    """float* of a C-contiguous float32 numpy array."""
    return a.ctypes.data_as(C.POINTER(C.c_float))
