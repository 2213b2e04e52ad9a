"""ctypes prototypes of every entry point libg6.so and libg4x.so export.

``GRAPE6`` lists the entry points of include/grape6.h (the reference ABI,
include/grape6.h:18-102 of the Starlab tree); ``G6X`` those of include/g6x.h;
``G6E`` those of include/g6ens.h (the ensemble engine).  ``GRAPE4`` and
``G4X`` are those of libg4x.so: include/grape4.h (reference
include/grape4.h:22-45) and include/g4x.h.
"""
from __future__ import annotations

import ctypes as C

_i = C.c_int
_pi = C.POINTER(C.c_int)
_pd = C.POINTER(C.c_double)
_vp = C.c_void_p

# name: (restype, [argtypes])
GRAPE6 = {
    "g6_set_tunit_": (None, [_pi]),
    "g6_set_xunit_": (None, [_pi]),
    "g6_open_": (_i, [_pi]),
    "g6_close_": (_i, [_pi]),
    "g6_open_special": (_i, [_i, _pi]),
    "g6_open_special2": (_i, [_i, _pi, _i]),
    "g6_reset_": (_i, [_pi]),
    "g6_reset_fofpga_": (_i, [_pi]),
    "g6_npipes_": (_i, []),
    "g6_set_ti_": (None, [_pi, _pd]),
    "g6_set_j_particle_": (_i, [_pi, _pi, _pi, _pd, _pd, _pd, _vp, _vp, _vp, _vp, _vp]),
    "g6calc_firsthalf_": (None, [_pi, _pi, _pi, _vp, _vp, _vp, _vp, _vp, _vp, _pd, _vp]),
    "g6calc_lasthalf_": (_i, [_pi, _pi, _pi, _vp, _vp, _vp, _pd, _vp, _vp, _vp, _vp]),
    "g6calc_lasthalf2_": (_i, [_pi, _pi, _pi, _vp, _vp, _vp, _pd, _vp, _vp, _vp, _vp, _vp]),
    "g6_read_neighbour_list_": (_i, [_pi]),
    "g6_read_neighbour_list_old_": (_i, [_pi]),
    "g6_get_neighbour_list_": (_i, [_pi, _pi, _pi, _pi, _vp]),
    "g6_initialize_jp_buffer_": (_i, [_pi, _pi]),
    "g6_flush_jp_buffer_": (_i, [_pi]),
}

G6X = {
    "g6x_set_j_particles": (_i, [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "g6x_self_sweep": (_i, [_i, _i, _i, _i, C.c_double, C.c_double]),
    "g6x_self_sweep_noalloc": (_i, [_i, _i, _i, _i, C.c_double, C.c_double, _vp, _vp, _vp,
                                    _vp]),
    "g6x_potential": (_i, [_i, _i, _i, _i, C.c_double, C.c_double, _vp]),
    "g6x_densities": (_i, [_i, _i, _i, _i, C.c_double, _i, _vp, _vp]),
    "g6x_read_results": (_i, [_i, _i, _vp, _vp, _vp, _vp]),
    "g6x_read_predicted": (_i, [_i, _i, _vp, _vp]),
    "g6x_sync": (_i, [_i]),
    "g6x_set_option": (_i, [C.c_char_p, C.c_longlong]),
    "g6x_get_option": (C.c_longlong, [C.c_char_p]),
    "g6x_kernel_time": (_i, [_i, _pd]),
    "g6x_reset_timing": (None, []),
    "g6x_last_force_geometry": (_i, [_pi, _pi, _pi]),
    "g6x_last_refinements": (_i, []),
    "g6x_autotune": (_i, [_i, _i, _i, C.c_double]),
    "g6x_get_tuning": (_i, [_i, _pi, _pi, _pi, _pi]),
    "g6x_clear_tuning": (_i, []),
    "g6x_save_tuning": (_i, [C.c_char_p]),
    "g6x_load_tuning": (_i, [C.c_char_p]),
    "g6x_last_error": (C.c_char_p, []),
}


G6E = {
    "g6e_open": (_i, [C.c_longlong, _i]),
    "g6e_close": (_i, []),
    "g6e_submit": (C.c_longlong, [_i, _vp, _vp, _vp, C.c_double, C.c_double, C.c_double]),
    "g6e_run": (_i, [_i]),
    "g6e_poll": (_i, [_i, _vp]),
    "g6e_result": (_i, [C.c_longlong, _i, _vp, _vp, _pd, C.POINTER(C.c_longlong), _pd, _pd]),
    "g6e_stats": (_i, [C.POINTER(C.c_longlong), _pd, C.POINTER(C.c_longlong), _pd]),
    "g6e_set_option": (_i, [C.c_char_p, C.c_longlong]),
    "g6e_get_option": (C.c_longlong, [C.c_char_p]),
    "g6e_last_error": (C.c_char_p, []),
}


# libg4x.so: the GRAPE-4 ABI (include/grape4.h; reference include/grape4.h:22-45)
GRAPE4 = {
    "h3open_": (None, []),
    "h3close_": (None, []),
    "h3npipe_": (_i, []),
    "h3setnboards_": (None, [_pi]),
    "h3getnboards_": (_i, []),
    "h3wait_": (C.c_uint, []),
    "h3setti_": (None, [_pd]),
    "h3setmode_": (None, [_pi, _pi]),
    "h3setled_": (None, [_pi, _pi]),
    "h3jpmax_": (_i, []),
    "h3jpdma_indirect_": (None, [_pi, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _pi]),
    "h3mjpdma_indirect_": (None, [_pi, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _pi, _pi]),
    "h3mjpdma_start_": (None, [_pi]),
    "h3mjpdma_flush_": (None, []),
    "h3calc_": (None, [_pi, _pi, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "h3calc_firsthalf_": (None, [_pi, _pi, _vp, _vp, _vp, _vp]),
    "h3calc_lasthalf_": (None, [_pi, _vp, _vp, _vp]),
    "h3nbread_": (None, [_pi]),
    "h3nblist_": (_i, [_pi, _pi, _vp]),
    "count_nblist_low": (_i, [_i, _i]),
    "h3setdebuglevel_": (None, [_pi]),
    "h3checkdmabuffers_": (None, []),
    "set_time_check_mode": (None, [_i]),
}

G4X = {
    "g4x_errors": (C.c_longlong, []),
    "g4x_last_error": (C.c_char_p, []),
    "g4x_set_option": (_i, [C.c_char_p, C.c_longlong]),
    "g4x_get_option": (C.c_longlong, [C.c_char_p]),
}


def _apply(lib: C.CDLL, tables) -> None:
    for table in tables:
        for name, (res, args) in table.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args


def declare(lib: C.CDLL) -> None:
    _apply(lib, (GRAPE6, G6X, G6E))


def declare_g4(lib: C.CDLL) -> None:
    _apply(lib, (GRAPE4, G4X))
