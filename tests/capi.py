"""ctypes binding of include/shwfs.h for the tests (the binding an FFI user writes)."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "compass-saxoplus_amd", "lib", "libshwfs.so")
HEADER = os.path.join(ROOT, "include", "shwfs.h")

SHWFS_OK = 0
SHWFS_ERR_ARG = -1
SHWFS_ERR_UNSUPPORTED = -2
SHWFS_ERR_HIP = -3
SHWFS_ERR_STATE = -4

METHOD_FFT_FP32, METHOD_DFT_BF16X3, METHOD_DFT_FP16 = 0, 1, 2
CENTROID_NONE, CENTROID_COG, CENTROID_TCOG, CENTROID_WCOG, CENTROID_MLP = 0, 1, 2, 3, 4

i32, u32, u64, f32, vp = C.c_int32, C.c_uint32, C.c_uint64, C.c_float, C.c_void_p


class ShGeom(C.Structure):
    _fields_ = [("nvalid", i32), ("nphase", i32), ("npup", i32), ("nfft", i32), ("ntot", i32),
                ("npix", i32), ("nrebin", i32), ("nxsub", i32),
                ("d_istart", vp), ("d_jstart", vp), ("d_validsubsx", vp), ("d_validsubsy", vp),
                ("d_offsets", vp), ("d_fluxPerSub", vp), ("d_binmap", vp),
                ("binmap_block", i32), ("bin_origin", i32), ("d_hrmap", vp),
                ("d_ftkernel", vp), ("kernconv", i32), ("d_ftlgskern", vp),
                ("d_bgpix", vp), ("nbgpix", i32)]


class CentroidArgs(C.Structure):
    _fields_ = [("kind", i32), ("nvalid", i32), ("npix", i32), ("size", i32),
                ("d_validx", vp), ("d_validy", vp), ("offset", f32), ("scale", f32),
                ("threshold", f32), ("d_weights", vp), ("d_ref", vp), ("d_slopes", vp),
                ("d_intensities", vp), ("slope_order", i32), ("fp64_moments", i32),
                ("d_mlp_params", vp), ("mlp_x_scale", f32), ("mlp_y_scale", f32),
                ("mlp_flux_lo", f32), ("mlp_flux_hi", f32), ("d_mlp_fallback", vp)]


class Layer(C.Structure):
    _fields_ = [("d_screen", vp), ("dim", i32), ("xoff", f32), ("yoff", f32), ("G", f32),
                ("thetaML", f32), ("dx", f32), ("dy", f32), ("delta", f32)]


class FrameArgs(C.Structure):
    _fields_ = [("d_phase", vp), ("d_pupil", vp), ("scale", f32), ("nphot", f32), ("noise", f32),
                ("noise_background", f32), ("seed", u64), ("call", u32), ("method", i32),
                ("wg_size", i32), ("d_bincube", vp), ("d_binimg", vp), ("d_intensities", vp),
                ("centroid", C.POINTER(CentroidArgs)), ("d_pupil_unit", vp),
                ("layers", C.POINTER(Layer)), ("nlayers", i32)]


class BatchTile(C.Structure):
    _fields_ = [("method", i32), ("split", i32)]


def declared_functions():
    """Names of every function declared in include/shwfs.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int32_t|const char\*)\s+(shwfs_\w+)\s*\(", txt, re.M)))


def load():
    lib = C.CDLL(LIB_PATH)
    lib.shwfs_sh_comp_image.argtypes = [C.POINTER(ShGeom), C.POINTER(FrameArgs), vp]
    lib.shwfs_sh_comp_image_noalloc.argtypes = [C.POINTER(ShGeom), C.POINTER(FrameArgs), vp]
    lib.shwfs_sh_check_geom.argtypes = [C.POINTER(ShGeom), i32]
    lib.shwfs_sh_unit_pupil.argtypes = [C.POINTER(ShGeom), vp, vp, vp]
    lib.shwfs_binmap_block_origin.argtypes = [vp, i32, i32, i32, C.POINTER(i32)]
    lib.shwfs_centroids.argtypes = [vp, vp, C.POINTER(CentroidArgs), vp]
    lib.shwfs_fill_weights.argtypes = [vp, vp, i32, i32, vp]
    lib.shwfs_prng.argtypes = [vp, C.c_int64, C.c_char, f32, f32, u64, u32, u32, vp]
    lib.shwfs_calibrate.argtypes = [vp, vp, vp, vp, vp, C.c_int64, vp]
    lib.shwfs_calibrate_u16.argtypes = [vp, vp, vp, vp, vp, C.c_int64, vp]
    lib.shwfs_digitalize.argtypes = [vp, vp, vp, vp, i32, i32, C.c_int64, vp]
    lib.shwfs_centroids_bpcog.argtypes = [vp, i32, C.POINTER(CentroidArgs), vp]
    lib.shwfs_field_stop.argtypes = [vp, vp, i32, f32, vp, i32, vp, vp, vp, vp]
    lib.shwfs_centroids_corr.argtypes = [vp, i32, vp, vp, i32, i32, vp, i32, vp, vp, i32, i32, f32,
                                         f32, vp, i32, vp, vp, vp, vp]
    lib.shwfs_fill_bincube.argtypes = [vp, vp, vp, vp, i32, i32, i32, vp]
    lib.shwfs_fill_bincube_2d.argtypes = [vp, vp, vp, i32, i32, i32, vp]
    lib.shwfs_slopes_geom.argtypes = [i32, vp, vp, vp, vp, vp, vp, i32, i32, f32, vp]
    lib.shwfs_tt_filter.argtypes = [vp, vp, vp, vp, i32, vp]
    lib.shwfs_fill_binimg.argtypes = [vp, vp, vp, vp, i32, i32, i32, f32, u64, u32, u32, vp]
    lib.shwfs_upload.argtypes = [vp, C.c_int64, vp, C.c_int64, i32, vp]
    lib.shwfs_raytrace.argtypes = [vp, i32, i32, C.POINTER(Layer), i32, i32, vp]
    lib.shwfs_mfma_probe.argtypes = [vp, vp, vp, i32, vp]
    lib.shwfs_batch_create.argtypes = [C.POINTER(ShGeom), C.POINTER(FrameArgs), i32,
                                       C.POINTER(vp)]
    lib.shwfs_batch_destroy.argtypes = [vp]
    lib.shwfs_batch_set_tiles.argtypes = [vp, C.POINTER(BatchTile), i32, i32]
    lib.shwfs_batch_get_tiles.argtypes = [vp, C.POINTER(BatchTile), C.POINTER(i32), C.POINTER(i32)]
    lib.shwfs_batch_set_frame.argtypes = [vp, u32, vp]
    lib.shwfs_batch_get_frame.argtypes = [vp, C.POINTER(u32)]
    lib.shwfs_batch_launch.argtypes = [vp, vp]
    lib.shwfs_batch_capture.argtypes = [vp, i32, vp]
    lib.shwfs_batch_replay.argtypes = [vp, i32, vp]
    lib.shwfs_batch_autotune.argtypes = [vp, u32, i32, vp, C.POINTER(C.c_float)]
    lib.shwfs_last_error.argtypes = [C.c_char_p, C.c_size_t]
    lib.shwfs_version.restype = C.c_char_p
    return lib


def last_error(lib):
    buf = C.create_string_buffer(512)
    lib.shwfs_last_error(buf, 512)
    return buf.value.decode()


def ptr(t):
    """Device address of a torch tensor (or None)."""
    return None if t is None else C.c_void_p(t.data_ptr())
