"""ctypes declarations for the C ABI in include/libfive.h.

Mirrors the role of the reference's Python binding module
(libfive/bind/python/libfive/ffi.py:141-154, which declares eval_f/r/d and
render_mesh): load the shared library and give every entry point its C
signature.  The library is the in-tree build ``libfive_amd/libfive_amd.so``;
importing this module never builds or falls back to anything else.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfive_amd.so")


class libfive_interval(ctypes.Structure):
    _fields_ = [("lower", ctypes.c_float), ("upper", ctypes.c_float)]


class libfive_region3(ctypes.Structure):
    _fields_ = [("X", libfive_interval), ("Y", libfive_interval),
                ("Z", libfive_interval)]


class libfive_region2(ctypes.Structure):
    _fields_ = [("X", libfive_interval), ("Y", libfive_interval)]


class libfive_vec2(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float)]


class libfive_contour(ctypes.Structure):
    _fields_ = [("pts", ctypes.POINTER(libfive_vec2)), ("count", ctypes.c_uint32)]


class libfive_contours(ctypes.Structure):
    _fields_ = [("cs", ctypes.POINTER(libfive_contour)), ("count", ctypes.c_uint32)]


class libfive_slice_segments(ctypes.Structure):
    _fields_ = [("verts", ctypes.POINTER(ctypes.c_float)),
                ("keys", ctypes.POINTER(ctypes.c_uint32)),
                ("vert_count", ctypes.c_uint32),
                ("segs", ctypes.POINTER(ctypes.c_uint32)),
                ("seg_count", ctypes.c_uint32),
                ("cells", ctypes.POINTER(ctypes.c_uint32)),
                ("cell_count", ctypes.c_uint32)]


class libfive_pixels(ctypes.Structure):
    _fields_ = [("pixels", ctypes.POINTER(ctypes.c_bool)),
                ("width", ctypes.c_uint32),
                ("height", ctypes.c_uint32)]


class libfive_vec3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float),
                ("z", ctypes.c_float)]


class libfive_tri(ctypes.Structure):
    _fields_ = [("a", ctypes.c_uint32), ("b", ctypes.c_uint32),
                ("c", ctypes.c_uint32)]


class libfive_mesh_coords(ctypes.Structure):
    _fields_ = [("verts", ctypes.POINTER(libfive_vec3)),
                ("vert_count", ctypes.c_uint32),
                ("coord_indices", ctypes.POINTER(ctypes.c_int32)),
                ("coord_index_count", ctypes.c_uint32)]


class libfive_vars(ctypes.Structure):
    _fields_ = [("vars", ctypes.POINTER(ctypes.c_void_p)),
                ("values", ctypes.POINTER(ctypes.c_float)),
                ("size", ctypes.c_uint32)]


class libfive_mesh(ctypes.Structure):
    _fields_ = [("verts", ctypes.POINTER(libfive_vec3)),
                ("tris", ctypes.POINTER(libfive_tri)),
                ("tri_count", ctypes.c_uint32),
                ("vert_count", ctypes.c_uint32)]


class libfive_mesh_settings(ctypes.Structure):
    _fields_ = [("min_feature", ctypes.c_double),
                ("max_err", ctypes.c_double),
                ("precision", ctypes.c_uint32),
                ("collect_stats", ctypes.c_uint32),
                ("cancel", ctypes.POINTER(ctypes.c_int))]


tree_t = ctypes.c_void_p
tape_t = ctypes.c_void_p
fp = ctypes.POINTER(ctypes.c_float)
u32p = ctypes.POINTER(ctypes.c_uint32)
u8p = ctypes.POINTER(ctypes.c_uint8)
szp = ctypes.POINTER(ctypes.c_size_t)

# name: (restype, argtypes)
SIGNATURES = {
    "libfive_mesh_delete": (None, [ctypes.POINTER(libfive_mesh)]),
    "libfive_opcode_enum": (ctypes.c_int, [ctypes.c_char_p]),
    "libfive_opcode_args": (ctypes.c_int, [ctypes.c_int]),
    "libfive_tree_x": (tree_t, []),
    "libfive_tree_y": (tree_t, []),
    "libfive_tree_z": (tree_t, []),
    "libfive_tree_var": (tree_t, []),
    "libfive_tree_is_var": (ctypes.c_bool, [tree_t]),
    "libfive_tree_const": (tree_t, [ctypes.c_float]),
    "libfive_tree_get_const": (ctypes.c_float,
                               [tree_t, ctypes.POINTER(ctypes.c_bool)]),
    "libfive_tree_nullary": (tree_t, [ctypes.c_int]),
    "libfive_tree_unary": (tree_t, [ctypes.c_int, tree_t]),
    "libfive_tree_binary": (tree_t, [ctypes.c_int, tree_t, tree_t]),
    "libfive_tree_id": (ctypes.c_void_p, [tree_t]),
    "libfive_tree_delete": (None, [tree_t]),
    "libfive_tree_remap": (tree_t, [tree_t, tree_t, tree_t, tree_t]),
    "libfive_tree_optimized": (tree_t, [tree_t]),
    "libfive_tree_print": (ctypes.c_void_p, [tree_t]),
    "libfive_free_str": (None, [ctypes.c_void_p]),
    "libfive_tree_eval_f": (ctypes.c_float, [tree_t, libfive_vec3]),
    "libfive_tree_eval_r": (libfive_interval, [tree_t, libfive_region3]),
    "libfive_tree_eval_d": (libfive_vec3, [tree_t, libfive_vec3]),
    "libfive_tree_render_mesh": (ctypes.POINTER(libfive_mesh),
                                 [tree_t, libfive_region3, ctypes.c_float]),
    "libfive_tree_render_mesh_st": (ctypes.POINTER(libfive_mesh),
                                    [tree_t, libfive_region3, ctypes.c_float]),
    "libfive_last_error": (ctypes.c_char_p, []),
    "libfive_gpu_available": (ctypes.c_int, []),
    "libfive_tree_eval_array": (ctypes.c_int,
                                [tree_t, fp, fp, fp, ctypes.c_size_t, fp]),
    "libfive_tree_eval_deriv_array": (ctypes.c_int,
                                      [tree_t, fp, fp, fp, ctypes.c_size_t,
                                       fp, fp, fp, fp]),
    "libfive_tree_eval_interval_array": (ctypes.c_int,
                                         [tree_t, fp, fp, fp, fp, fp, fp,
                                          ctypes.c_size_t, fp, fp, u8p]),
    "libfive_tree_eval_ambiguous_array": (ctypes.c_int,
                                          [tree_t, fp, fp, fp,
                                           ctypes.c_size_t, u8p]),
    "libfive_tree_eval_deriv_ambiguous_array": (ctypes.c_int,
                                                [tree_t, fp, fp, fp,
                                                 ctypes.c_size_t, u8p]),
    "libfive_tape_new": (tape_t, [tree_t]),
    "libfive_tape_delete": (None, [tape_t]),
    "libfive_tape_clauses": (ctypes.c_uint32, [tape_t]),
    "libfive_tape_registers": (ctypes.c_uint32, [tape_t]),
    "libfive_tape_eval_values_device": (ctypes.c_int,
                                        [tape_t, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_size_t, ctypes.c_void_p,
                                         ctypes.c_void_p]),
    "libfive_tape_eval_derivs_device": (ctypes.c_int,
                                        [tape_t, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_size_t, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p]),
    "libfive_tape_eval_values_noalloc": (ctypes.c_int, [tape_t, fp, fp, fp,
                                                        ctypes.c_size_t, fp]),
    "libfive_tape_eval_values": (ctypes.c_int, [tape_t, fp, fp, fp,
                                                ctypes.c_size_t, fp]),
    "libfive_tape_eval_derivs": (ctypes.c_int, [tape_t, fp, fp, fp,
                                                ctypes.c_size_t, fp, fp, fp,
                                                fp]),
    "libfive_trees_render_mesh_batch": (
        ctypes.c_int, [ctypes.POINTER(tree_t), ctypes.POINTER(libfive_region3),
                       fp, ctypes.c_size_t, ctypes.c_int,
                       ctypes.POINTER(ctypes.POINTER(libfive_mesh))]),
    "libfive_trees_render_mesh_merged": (
        ctypes.c_int, [ctypes.POINTER(tree_t), ctypes.POINTER(libfive_region3),
                       fp, ctypes.c_size_t,
                       ctypes.POINTER(ctypes.POINTER(libfive_mesh))]),
    "libfive_tree_render_mesh_coords": (ctypes.POINTER(libfive_mesh_coords),
                                        [tree_t, libfive_region3, ctypes.c_float]),
    "libfive_mesh_coords_delete": (None, [ctypes.POINTER(libfive_mesh_coords)]),
    "libfive_tree_render_pixels": (ctypes.POINTER(libfive_pixels),
                                   [tree_t, libfive_region2, ctypes.c_float,
                                    ctypes.c_float]),
    "libfive_pixels_delete": (None, [ctypes.POINTER(libfive_pixels)]),
    "libfive_tree_render_slice": (ctypes.POINTER(libfive_contours),
                                  [tree_t, libfive_region2, ctypes.c_float,
                                   ctypes.c_float]),
    "libfive_tree_render_slice3": (ctypes.c_void_p, [tree_t, libfive_region2,
                                                     ctypes.c_float, ctypes.c_float]),
    "libfive_contours_delete": (None, [ctypes.POINTER(libfive_contours)]),
    "libfive_contours3_delete": (None, [ctypes.c_void_p]),
    "libfive_tree_render_slice_segments": (
        ctypes.POINTER(libfive_slice_segments),
        [tree_t, libfive_region2, ctypes.c_float, ctypes.c_double, ctypes.c_double]),
    "libfive_slice_segments_delete": (None, [ctypes.POINTER(libfive_slice_segments)]),
    "libfive_tree_save_mesh": (ctypes.c_bool, [tree_t, libfive_region3,
                                               ctypes.c_float, ctypes.c_char_p]),
    "libfive_tree_save_meshes": (ctypes.c_bool, [ctypes.POINTER(tree_t),
                                                 libfive_region3, ctypes.c_float,
                                                 ctypes.c_float, ctypes.c_char_p]),
    "libfive_tree_evaluator": (ctypes.c_void_p, [tree_t, libfive_vars]),
    "libfive_evaluator_update_vars": (ctypes.c_bool, [ctypes.c_void_p, libfive_vars]),
    "libfive_evaluator_save_mesh": (ctypes.c_bool, [ctypes.c_void_p, libfive_region3,
                                                    ctypes.c_char_p]),
    "libfive_evaluator_delete": (None, [ctypes.c_void_p]),
    "libfive_vars_delete": (None, [ctypes.POINTER(libfive_vars)]),
    "libfive_tape_specialize": (ctypes.c_int,
                                [tape_t, ctypes.POINTER(ctypes.c_double)]),
    "libfive_tape_use_specialized": (None, [tape_t, ctypes.c_int]),
    "libfive_tree_specialized_code_size": (ctypes.c_size_t, [tree_t]),
    "libfive_set_points_per_lane": (None, [ctypes.c_int]),
    "libfive_tree_tape_info": (ctypes.c_int, [tree_t, u32p, u32p]),
    "libfive_autotune_enable": (None, [ctypes.c_int]),
    "libfive_autotune_reset": (None, []),
    "libfive_autotune_force": (None, [ctypes.c_char_p, ctypes.c_int]),
    "libfive_autotune_report": (ctypes.c_size_t, [ctypes.c_char_p, ctypes.c_size_t]),
    "libfive_tree_deck_vars": (ctypes.c_int, [tree_t, ctypes.POINTER(ctypes.c_uint32),
                                              ctypes.POINTER(ctypes.c_void_p),
                                              ctypes.POINTER(ctypes.c_size_t)]),
    "libfive_tree_deck": (ctypes.c_int, [tree_t, u32p, szp, u32p, fp, szp,
                                         u32p, u32p]),
    "libfive_mesh_vertex_keys": (u32p, [ctypes.POINTER(libfive_mesh)]),
    "libfive_last_mesh_level_counts": (ctypes.c_size_t,
                                       [ctypes.POINTER(ctypes.c_uint64),
                                        ctypes.c_size_t]),
    "libfive_last_mesh_tape_lengths": (ctypes.c_size_t, [u32p, ctypes.c_size_t]),
    "libfive_last_mesh_times": (None, [ctypes.POINTER(ctypes.c_double)]),
    "libfive_last_mesh_cells": (ctypes.c_size_t, [u32p, ctypes.c_size_t]),
    "libfive_mesh_settings_default": (None,
                                      [ctypes.POINTER(libfive_mesh_settings)]),
    "libfive_tree_render_mesh_settings": (
        ctypes.POINTER(libfive_mesh),
        [tree_t, libfive_region3, ctypes.POINTER(libfive_mesh_settings)]),
}

_lib = None


def lib():
    """Loads libfive_amd.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().libfive_last_error().decode()
