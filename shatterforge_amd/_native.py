"""ctypes binding of libsf_hip.so (the C ABI declared in include/sf.h).

The library is built in-tree (``shatterforge_amd/libsf_hip.so``); importing this module without it raises, so no
code path can silently fall back to anything else.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsf_hip.so")

c_int, c_uint, c_float, c_double, c_void_p, c_char_p = (ctypes.c_int, ctypes.c_uint32, ctypes.c_float, ctypes.c_double,
                                                        ctypes.c_void_p, ctypes.c_char_p)
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_int64 = ctypes.c_int64

# name: (restype, argtypes)
_PROTOS = {
    "sf_last_error": (c_int, []),
    "sf_last_error_message": (c_char_p, []),
    "sf_device_count": (c_int, []),
    "sf_world_create": (c_void_p, [c_int, c_int, c_void_p]),
    "sf_world_destroy": (None, [c_void_p]),
    "sf_world_set_params": (c_int, [c_void_p, c_void_p, c_int]),
    "sf_world_load": (c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p]),
    "sf_world_step": (c_int, [c_void_p, c_float]),
    "sf_world_step_many": (c_int, [c_void_p, c_float, c_int]),
    "sf_world_sync": (c_int, [c_void_p]),
    "sf_world_read_bodies": (c_int, [c_void_p, c_void_p]),
    "sf_world_stats": (c_int, [c_void_p, c_void_p]),
    "sf_world_nbody": (c_int, [c_void_p]),
    "sf_world_candidates": (c_int, [c_void_p, c_void_p, c_int]),
    "sf_world_manifolds": (c_int, [c_void_p, c_void_p, c_void_p, c_int]),
    "sf_world_colours": (c_int, [c_void_p, c_void_p, c_int]),
    "sf_world_islands": (c_int, [c_void_p, c_void_p]),
    "sf_world_pose": (ctypes.POINTER(c_float), [c_void_p]),
    "sf_world_particles_capacity": (c_int, [c_void_p, c_int]),
    "sf_world_add_particles": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "sf_world_kill_particle": (c_int, [c_void_p, c_int]),
    "sf_world_blast_particles": (c_int, [c_void_p, c_float, c_float, c_float, c_float, c_float]),
    "sf_world_wake_particles": (c_int, [c_void_p]),
    "sf_world_read_particles": (c_int, [c_void_p, c_void_p]),
    "sf_world_particle_stats": (c_int, [c_void_p, c_void_p]),
    "sf_world_particle_ms": (c_int, [c_void_p, c_void_p]),
    "sf_world_set_timing": (c_int, [c_void_p, c_int]),
    "sf_world_stage_ms": (c_int, [c_void_p, c_void_p]),
    "sf_world_set_kernel_timing": (c_int, [c_void_p, c_int]),
    "sf_world_kernel_stats": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "sf_world_set_graph": (c_int, [c_void_p, c_int]),
    "sf_world_graph_stats": (c_int, [c_void_p, c_void_p, c_void_p]),
    "sf_pile_setup": (None, [c_uint, c_uint]),
    "sf_stack_setup": (None, [c_uint, c_uint, c_uint]),
    "sf_step": (c_uint, [c_float]),
    "sf_body_count": (c_uint, []),
    "sf_manifolds": (c_uint, []),
    "sf_points": (c_uint, []),
    "sf_max_speed": (c_float, []),
    "sf_body_pos": (c_float, [c_uint, c_uint]),
    "sf_pose": (ctypes.POINTER(c_float), []),
    "sf_scene_pile": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "sf_build_body": (c_int, [c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "sf_hull_count": (c_int, []),
    "sf_hull_pattern_count": (c_int, []),
    "sf_hull_pattern": (c_int, [c_int, c_void_p]),
    "sf_hull_cell": (c_int, [c_int, c_void_p]),
    "sf_hull_vertices": (c_int, [c_int, c_int, c_void_p]),
    "sf_hull_face": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "sf_hull_edges": (c_int, [c_int, c_void_p, c_void_p]),
    "sf_hull_tables": (c_int, [c_void_p] * 7),
    "sf_hull_shapes": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "sf_scene_stack": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "sf_scene_rubble": (c_int, [c_int, c_float, c_int, c_float, c_float, c_float, c_float, c_float, c_uint,
                                c_void_p, c_void_p, c_void_p, c_void_p]),
    "sf_build_compound": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "sf_load_create": (c_void_p, []),
    "sf_load_destroy": (None, [c_void_p]),
    "sf_load_add": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "sf_load_solve": (c_int, [c_void_p, c_float, c_double, c_int]),
    "sf_load_solve_begin": (c_int, [c_void_p, c_float, c_double, c_int]),
    "sf_load_solve_ready": (c_int, [c_void_p]),
    "sf_load_solve_end": (c_int, [c_void_p]),
    "sf_load_info": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "sf_load_field": (c_int, [c_void_p, c_int, c_void_p]),
    "sf_load_verdicts": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "sf_load_topology": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "sf_load_count": (c_int, [c_void_p]),
    "sf_load_set_packed": (c_int, [c_void_p, c_int]),
    "sf_load_solve_static": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_double, c_int,
                                     c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "sf_load_static_capacity": (c_int64, []),
    "sf_load_last_timing": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "sf_scene_building": (c_int, [c_int, c_int, c_int, c_int, c_int, c_uint, c_void_p, c_void_p, c_void_p]),
    "sf_scene_tower": (c_int, [c_void_p, c_void_p, c_void_p]),
    "sf_scene_site": (c_int, [c_uint, c_int, c_void_p, c_void_p, c_void_p]),
    "sf_lat_create": (c_void_p, [c_int, c_int, c_int, c_int, c_int]),
    "sf_lat_destroy": (None, [c_void_p]),
    "sf_lat_scene_add": (c_int, [c_void_p, c_int, c_int, c_int, c_uint]),
    "sf_lat_scene_clear": (c_int, [c_void_p, c_int, c_uint]),
    "sf_lat_scene_release": (c_int, [c_void_p, c_int]),
    "sf_lat_scene_place": (c_int, [c_void_p, c_int, c_int, c_int, c_uint]),
    "sf_lat_bytes_resident": (c_int64, [c_void_p, c_int]),
    "sf_lat_bytes_for": (c_int64, [c_void_p, c_int, c_int, c_int]),
    "sf_serve_create": (c_void_p, [c_int, c_int, c_int, c_int, c_int64, c_int]),
    "sf_serve_destroy": (None, [c_void_p]),
    "sf_serve_query_bytes": (c_int64, [c_void_p, c_int]),
    "sf_serve_submit": (c_int, [c_void_p, ctypes.c_uint64, c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_uint]),
    "sf_serve_tick": (c_int, [c_void_p, c_float]),
    "sf_serve_poll": (c_int, [c_void_p, c_void_p, c_int]),
    "sf_serve_status": (c_int, [c_void_p, c_void_p]),
    "sf_serve_timing": (c_int, [c_void_p, c_void_p]),
    "sf_serve_world": (c_void_p, [c_void_p]),
    "sf_lat_add_structure": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "sf_lat_add_body": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "sf_lat_finalize": (c_int, [c_void_p, c_int]),
    "sf_lat_warm_all": (c_int, [c_void_p, c_int]),
    "sf_lat_wake_all": (c_int, [c_void_p, c_int]),
    "sf_lat_explode": (c_int, [c_void_p, c_int, c_float, c_float, c_float, c_float, c_float]),
    "sf_lat_step": (c_int, [c_void_p, c_float]),
    "sf_lat_stats": (c_int, [c_void_p, c_int, c_void_p]),
    "sf_lat_counts": (c_int, [c_void_p, c_int, c_void_p]),
    "sf_lat_bodies": (c_int, [c_void_p, c_int, c_void_p]),
    "sf_lat_bricks": (c_int, [c_void_p, c_int, c_void_p]),
    "sf_lat_structs": (c_int, [c_void_p, c_int, c_void_p]),
    "sf_lat_matstats": (c_int, [c_void_p, c_int, c_void_p]),
    "sf_lat_contacts": (c_int, [c_void_p, c_int, c_void_p, c_int]),
    "sf_lat_rng": (c_int, [c_void_p, c_int, c_void_p]),
    "sf_lat_set_field": (c_int, [c_void_p, c_int, c_int, c_void_p]),
    "sf_lat_failures": (c_int, [c_void_p, c_int, c_void_p]),
    "sf_lat_timing": (c_int, [c_void_p, c_void_p, c_void_p]),
    "sf_lat_phase_ms": (c_int, [c_void_p, c_void_p]),
    "sf_lat_set_fused": (c_int, [c_void_p, c_int]),
    "sf_lat_set_host_threads": (c_int, [c_void_p, c_int]),
    "sf_lat_set_conn_device": (c_int, [c_void_p, c_int]),
    "sf_lat_set_sor_lds": (c_int, [c_void_p, c_int]),
    "sf_legacy_init": (c_uint, []),
    "sf_legacy_step": (None, [c_float]),
    "sf_legacy_explode": (None, [c_float, c_float, c_float, c_float, c_float]),
    "sf_legacy_throw_charge": (c_uint, [c_float, c_float, c_float, c_float, c_float, c_float]),
    "sf_legacy_detonate": (c_uint, []),
    "sf_legacy_n_structs": (c_uint, []),
    "sf_legacy_ptr_structs": (c_void_p, []),
    "sf_legacy_ptr_stats": (c_void_p, []),
    "sf_legacy_clear_events": (None, []),
    "sf_legacy_n_events": (c_uint, []),
    "sf_legacy_ptr_events": (c_void_p, []),
    "sf_legacy_dbg_hist": (c_void_p, [c_uint]),
    "sf_legacy_dbg_bodies": (c_void_p, []),
    "sf_legacy_dbg_matstats": (c_void_p, []),
    "sf_legacy_dbg_speed": (c_void_p, []),
    "sf_legacy_dbg_bigbody": (c_void_p, []),
    "sf_legacy_dbg_top": (c_void_p, [c_uint]),
    "sf_legacy_dbg_wake_all": (None, []),
    "sf_bricks_create": (c_void_p, []),
    "sf_bricks_destroy": (None, [c_void_p]),
    "sf_bricks_len": (c_int, [c_void_p]),
    "sf_bricks_add": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int]),
    "sf_bricks_set_flags": (c_int, [c_void_p, c_int, c_int]),
    "sf_bricks_flags": (c_int, [c_void_p, c_int]),
    "sf_bricks_mass": (c_float, [c_void_p, c_int]),
    "sf_bricks_remove": (c_int, [c_void_p, c_int]),
    "sf_bricks_build_joints": (c_int, [c_void_p]),
    "sf_bricks_joints": (c_int, [c_void_p, c_void_p, c_int]),
    "sf_bricks_set_joint_damage": (c_int, [c_void_p, c_int, c_int]),
    "sf_bricks_save": (c_int64, [c_void_p, c_void_p, c_int64]),
    "sf_bricks_load": (c_void_p, [c_void_p, c_int64]),
    "sf_bricks_set_cell": (c_int, [c_void_p, c_int, c_int, c_int, c_int]),
    "sf_bricks_cell": (c_int, [c_void_p, c_int, c_void_p]),
    "sf_bricks_box": (c_int, [c_void_p, c_int, c_void_p]),
    "sf_warm_create": (c_void_p, []),
    "sf_warm_destroy": (None, [c_void_p]),
    "sf_warm_size": (c_int, [c_void_p]),
    "sf_net_create": (c_void_p, []),
    "sf_net_destroy": (None, [c_void_p]),
    "sf_net_add": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "sf_net_solve": (c_int, [c_void_p]),
    "sf_net_clear": (c_int, [c_void_p]),
    "sf_net_count": (c_int, [c_void_p]),
    "sf_net_set_host_threads": (c_int, [c_void_p, c_int]),
    "sf_net_set_route": (c_int, [c_void_p, c_int, c_int, c_int]),
    "sf_net_result": (c_int, [c_void_p, c_int, c_void_p]),
    "sf_net_levels": (c_int, [c_void_p, c_int, c_void_p]),
    "sf_net_joints": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int]),
    "sf_net_inner": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int]),
    "sf_net_cells": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int]),
    "sf_net_ms": (c_int, [c_void_p, c_void_p]),
    "sf_net_host_ms": (c_int, [c_void_p, c_void_p]),
    "sf_legacy_world": (c_void_p, [c_void_p]),
    "sf_lat_subphase_ms": (c_int, [c_void_p, c_void_p]),
}

_lib = None


def lib():
    """The loaded library (loaded once). Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class SfError(RuntimeError):
    pass


def check(rc):
    """Raise SfError with the library's message for a negative return code."""
    if rc is None or (isinstance(rc, int) and rc < 0):
        L = lib()
        raise SfError(f"libsf_hip error {L.sf_last_error()}: {L.sf_last_error_message().decode()}")
    return rc


def ptr(a):
    """Data pointer of a contiguous numpy array (or None)."""
    return None if a is None else a.ctypes.data_as(c_void_p)


def declared_symbols():
    """Function names declared in include/sf.h (the drop-in boundary)."""
    import re
    hdr = os.path.join(os.path.dirname(_HERE), "include", "sf.h")
    names = []
    with open(hdr) as f:
        for line in f:
            m = re.match(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(sf_[a-z_0-9]+)\s*\(", line)
            if m:
                names.append(m.group(1))
    return names
