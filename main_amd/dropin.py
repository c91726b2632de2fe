"""Per-call drop-ins of collective_code's module procedures on the GPU.

Each function has the name and argument meaning of the Fortran procedure it
replaces and updates its array arguments in place, as the Fortran intent(inout)
dummies do.  Integer arrays hold 1-based particle and cell numbers; 2-D arrays
are column-major with the reference's shapes.  Random draws are passed in
(the reference's grnd() calls of one thread, in call order), so a caller
keeps its Mersenne-Twister stream on the host.

    collective_code/mpc_rotation_mod.f95:6-36     mpc_rotation
    collective_code/reaction_mod.f95:8-27         reaction_simple
    collective_code/neighbour_list_mod.f95:6-26   neighbour_list
    collective_code/neighbour_list_mod.f95:63-87  inout_particle_list
    collective_code/distribute_mod.f95:6-40       distribute
    collective_code/sort_hilbert_mod.f95:6-24     sort_hilbert
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib, ptr, require_gpu

I32, I64, F64 = C.c_int32, C.c_int64, C.c_double


def _f64(a, name):
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and (a.flags.c_contiguous or a.flags.f_contiguous)):
        raise TypeError(f"{name} must be a contiguous float64 array")
    return a


def _i(a, dt, name):
    if not (isinstance(a, np.ndarray) and a.dtype == dt and (a.flags.c_contiguous or a.flags.f_contiguous)):
        raise TypeError(f"{name} must be a contiguous {np.dtype(dt).name} array")
    return a


def mpc_rotation(NT, rotang, numcell, a0, box, Xsol, Ysol, Zsol, VXsol, VYsol, VZsol, VXCM, VYCM, VZCM,
                 tCellShift, NewCellN, UVX, UVY, UVZ, draws):
    """Random shift, cell centre-of-mass velocities and rotation about a random
    axis by rotang (mpc_rotation_mod.f95:83-232).  draws: 3 + 2*numcell."""
    require_gpu()
    d = np.ascontiguousarray(draws, dtype=np.float64)
    if d.size < 3 + 2 * int(numcell):
        raise ValueError("mpc_rotation needs 3 + 2*numcell draws")
    for n, a in (("VXCM", VXCM), ("VYCM", VYCM), ("VZCM", VZCM), ("tCellShift", tCellShift), ("UVX", UVX),
                 ("UVY", UVY), ("UVZ", UVZ)):
        if _f64(a, n).size < numcell:
            raise ValueError(f"{n} shorter than numcell")
    for n, a in (("Xsol", Xsol), ("Ysol", Ysol), ("Zsol", Zsol), ("VXsol", VXsol), ("VYsol", VYsol),
                 ("VZsol", VZsol)):
        if _f64(a, n).size < NT:
            raise ValueError(f"{n} shorter than NT")
    if _i(NewCellN, np.int64, "NewCellN").size < NT:
        raise ValueError("NewCellN shorter than NT")
    b = np.ascontiguousarray(box, dtype=np.float64)
    check(lib().mpcx_mpc_rotation(I32(NT), F64(rotang), I64(numcell), F64(a0), ptr(b), ptr(Xsol), ptr(Ysol),
                                  ptr(Zsol), ptr(VXsol), ptr(VYsol), ptr(VZsol), ptr(VXCM), ptr(VYCM), ptr(VZCM),
                                  ptr(tCellShift), ptr(NewCellN), ptr(UVX), ptr(UVY), ptr(UVZ), ptr(d)),
          "mpc_rotation")


def reaction_simple(MPCt, numcell, k1, NT, flag, tBCellN, sBCellN, draws):
    """B -> A in cell i with probability 1 - exp(-k1*tBCellN(i)*MPCt), applied to
    the cell's first B member sBCellN(i,1) (reaction_mod.f95:41-62).
    sBCellN is the (vol, MaxB) column-major member array; draws: numcell."""
    require_gpu()
    _i(flag, np.int32, "flag")
    tb = _i(tBCellN, np.int64, "tBCellN")
    sb = _i(sBCellN, np.int64, "sBCellN")
    if not sb.flags.f_contiguous:
        raise TypeError("sBCellN must be column-major (vol, MaxB)")
    vol = sb.shape[0]
    d = np.ascontiguousarray(draws, dtype=np.float64)
    if d.size < numcell or tb.size < numcell or numcell > vol:
        raise ValueError("reaction_simple: arrays shorter than numcell")
    check(lib().mpcx_reaction_simple(F64(MPCt), I64(numcell), F64(k1), I32(NT), ptr(flag), ptr(tb), ptr(sb),
                                     I64(vol), ptr(d)), "reaction_simple")


def neighbour_list(NT, box, cutoff, skinrad, Xsol, Ysol, Zsol, Xdim, Ydim, Zdim, neighbour_num, neighbour):
    """Appends to ``neighbour`` every particle within cutoff+skinrad of the
    sphere at (Xdim, Ydim, Zdim), ascending (neighbour_list_mod.f95:34-57).
    neighbour_num: a one-element int64 array (the count, updated); neighbour:
    a 1-D int64 array or a strided view such as neighbour[k, i, :] of the
    reference's (numdims, 2, max_neighbour) column-major array."""
    require_gpu()
    if not (isinstance(neighbour, np.ndarray) and neighbour.dtype == np.int64 and neighbour.ndim == 1):
        raise TypeError("neighbour must be a 1-D int64 array or view")
    if neighbour.strides[0] <= 0 or neighbour.strides[0] % 8:
        raise ValueError("neighbour must have a positive element stride")
    nn = _i(neighbour_num, np.int64, "neighbour_num")
    b = np.ascontiguousarray(box, dtype=np.float64)
    stride = neighbour.strides[0] // 8
    base = neighbour.ctypes.data_as(C.c_void_p)
    check(lib().mpcx_neighbour_list(I32(NT), ptr(b), F64(cutoff), F64(skinrad), ptr(_f64(Xsol, "Xsol")),
                                    ptr(_f64(Ysol, "Ysol")), ptr(_f64(Zsol, "Zsol")), F64(Xdim), F64(Ydim),
                                    F64(Zdim), ptr(nn), base, I64(stride), I64(neighbour.shape[0])),
          "neighbour_list")


def inout_particle_list_noalloc(NT, box, cutoffC, cutoffN, skinrad, xsol_dev, ysol_dev, zsol_dev, xdim, ydim, zdim,
                                outside_particle_num, outside_particle_dev, outside_capacity, inside_particle_num,
                                inside_particle_dev, inside_capacity, numdims):
    """inout_particle_list on device arrays, with no device allocation: the
    solvent coordinates and both lists are device addresses (ints) the caller
    owns, with room for outside_capacity / inside_capacity entries; the counts
    (one-element int64 arrays), box and sphere arrays stay on the host.  At
    most 67,108,864 particles and 64 dimers; calls that overlap on one GPU
    race (they share the library's fixed scratch array)."""
    require_gpu()
    on = _i(outside_particle_num, np.int64, "outside_particle_num")
    inn = _i(inside_particle_num, np.int64, "inside_particle_num")
    dims = [np.asfortranarray(a, dtype=np.float64) for a in (xdim, ydim, zdim)]
    b = np.ascontiguousarray(box, dtype=np.float64)
    V = C.c_void_p
    check(lib().mpcx_inout_particle_list_noalloc(I32(NT), ptr(b), F64(cutoffC), F64(cutoffN), F64(skinrad),
                                                 V(xsol_dev), V(ysol_dev), V(zsol_dev), *[ptr(a) for a in dims],
                                                 ptr(on), V(outside_particle_dev), I64(outside_capacity), ptr(inn),
                                                 V(inside_particle_dev), I64(inside_capacity), I32(numdims)),
          "inout_particle_list_noalloc")


def inout_particle_list(NT, box, cutoffC, cutoffN, skinrad, xsol, ysol, zsol, xdim, ydim, zdim,
                        outside_particle_num, outside_particle, inside_particle_num, inside_particle, numdims):
    """Splits the solvent into particles within cutoff+skinrad of some sphere
    (inside) and the rest (outside), both ascending, appended to the counts
    (neighbour_list_mod.f95:105-177).  The counts are one-element int64 arrays;
    xdim/ydim/zdim are the (numdims, 2) sphere arrays."""
    require_gpu()
    on = _i(outside_particle_num, np.int64, "outside_particle_num")
    inn = _i(inside_particle_num, np.int64, "inside_particle_num")
    _i(outside_particle, np.int64, "outside_particle")
    _i(inside_particle, np.int64, "inside_particle")
    if outside_particle.size < on[0] + NT:
        raise ValueError("outside_particle must have room for NT more entries")
    dims = [np.asfortranarray(a, dtype=np.float64) for a in (xdim, ydim, zdim)]
    b = np.ascontiguousarray(box, dtype=np.float64)
    check(lib().mpcx_inout_particle_list(I32(NT), ptr(b), F64(cutoffC), F64(cutoffN), F64(skinrad),
                                         ptr(_f64(xsol, "xsol")), ptr(_f64(ysol, "ysol")), ptr(_f64(zsol, "zsol")),
                                         *[ptr(a) for a in dims], ptr(on), ptr(outside_particle), ptr(inn),
                                         ptr(inside_particle), I64(inside_particle.size), I32(numdims)),
          "inout_particle_list")


def distribute(NT, numdims, flag, Xsol, Ysol, Zsol, Xdim, Ydim, Zdim, box, tACellN, tBCellN, tSCellN,
               sACellN, sBCellN, sSCellN, cutoffC2, cutoffN2, outside_particle_num, outside_particle,
               inside_particle_num, inside_particle):
    """Per-cell census of A, B and S solvent for the reaction: the outside
    particles, then the inside ones beyond every sphere's LJ cutoff, in list
    order (distribute_mod.f95:81-201).  t?CellN: (vol,) int64 counts;
    s?CellN: (vol, maxs) column-major int64 member lists."""
    require_gpu()
    vol = tACellN.size
    for n, a in (("tACellN", tACellN), ("tBCellN", tBCellN), ("tSCellN", tSCellN)):
        if _i(a, np.int64, n).size != vol:
            raise ValueError("count arrays must have one entry per cell")
    maxs = None
    for n, a in (("sACellN", sACellN), ("sBCellN", sBCellN), ("sSCellN", sSCellN)):
        _i(a, np.int64, n)
        if not a.flags.f_contiguous or a.ndim != 2 or a.shape[0] != vol:
            raise TypeError(f"{n} must be a column-major (vol, maxs) array")
        maxs = a.shape[1] if maxs is None else min(maxs, a.shape[1])
    dims = [np.asfortranarray(a, dtype=np.float64) for a in (Xdim, Ydim, Zdim)]
    b = np.ascontiguousarray(box, dtype=np.float64)
    ol = np.ascontiguousarray(outside_particle, dtype=np.int64)
    il = np.ascontiguousarray(inside_particle, dtype=np.int64)
    check(lib().mpcx_distribute(I32(NT), I32(numdims), ptr(_i(flag, np.int32, "flag")), ptr(_f64(Xsol, "Xsol")),
                                ptr(_f64(Ysol, "Ysol")), ptr(_f64(Zsol, "Zsol")), *[ptr(a) for a in dims], ptr(b),
                                I64(vol), ptr(tACellN), ptr(tBCellN), ptr(tSCellN), ptr(sACellN), ptr(sBCellN),
                                ptr(sSCellN), I64(maxs), F64(cutoffC2), F64(cutoffN2), I64(outside_particle_num),
                                ptr(ol), I64(inside_particle_num), ptr(il)), "distribute")


def sort_hilbert(vol, NT, box, tour, xsol, ysol, zsol, flag, vxsol, vysol, vzsol):
    """Reorders the solvent arrays cell by cell along ``tour`` (vol 1-based
    cell numbers, every cell once), ascending particle index within a cell
    (sort_hilbert_mod.f95:47-92)."""
    require_gpu()
    t = np.ascontiguousarray(tour, dtype=np.int64)
    if t.size != vol:
        raise ValueError("tour must list vol cells")
    b = np.ascontiguousarray(box, dtype=np.float64)
    arrs = [_f64(a, n) for n, a in (("xsol", xsol), ("ysol", ysol), ("zsol", zsol))]
    vel = [_f64(a, n) for n, a in (("vxsol", vxsol), ("vysol", vysol), ("vzsol", vzsol))]
    check(lib().mpcx_sort_hilbert(I64(vol), I32(NT), ptr(b), ptr(t), *[ptr(a) for a in arrs],
                                  ptr(_i(flag, np.int32, "flag")), *[ptr(a) for a in vel]), "sort_hilbert")
