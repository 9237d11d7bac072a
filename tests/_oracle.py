"""ctypes bindings of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure only: parity tests, smoke() and bench.py's CPU-baseline
leg use it as the checker.  Nothing in the product imports this module.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

U64P = C.POINTER(C.c_uint64)
U32P = C.POINTER(C.c_uint32)
F32P = C.POINTER(C.c_float)
F64P = C.POINTER(C.c_double)


class Bulk(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("mutr", "mua", "g", "n", "flc", "muaf", "eY", "albedof")]


class Sim(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("esp", C.c_float), ("gs", C.c_float), ("halfx", C.c_float), ("halfy", C.c_float),
        ("n_bulks", C.c_uint32), ("bulks", C.POINTER(Bulk)), ("labels", C.POINTER(C.c_uint8)),
        ("det_dx", C.c_float), ("det_dy", C.c_float), ("det_x0", C.c_float), ("det_y0", C.c_float),
        ("det_nx", C.c_int32), ("det_ny", C.c_int32),
        ("do_temp", C.c_int32), ("t_ndx", C.c_uint32), ("t_ndy", C.c_uint32),
        ("t_posx", F32P), ("t_posy", F32P), ("t_r", C.c_float), ("t_rort", C.c_uint32),
        ("t_bins", C.c_uint64), ("t_max", C.c_uint64),
        ("src_kind", C.c_int32), ("xi", C.c_float), ("yi", C.c_float), ("zi", C.c_float), ("src_r", C.c_float),
        ("start_weight", C.c_uint32),
        ("fhd_on", C.c_int32), ("axyz_on", C.c_int32), ("sq_on", C.c_int32),
        ("layered", C.c_int32), ("z_min", F32P), ("z_max", F32P), ("inc", Bulk),
        ("inc_x", C.c_float), ("inc_y", C.c_float), ("inc_z", C.c_float), ("inc_r", C.c_float),
        ("start_layer", C.c_uint32),
    ]


NTOT = 16
TOT_NAMES = ["launched_w", "R", "T", "A", "roul_gain", "roul_loss", "zero_kill", "photons", "steps", "aborted"]


class Tallies(C.Structure):
    _fields_ = [("Rd", U64P), ("Tt", U64P), ("fhd", U64P), ("axyz", U64P), ("time", U64P),
                ("Rd2", U64P), ("Tt2", U64P), ("tot", C.c_uint64 * NTOT)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.orc_init_rng.argtypes = [U64P, U32P, C.c_uint32, C.c_char_p, C.c_uint64]
        L.orc_run_photons.argtypes = [C.POINTER(Sim), U32P, C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint64,
                                      C.c_double, C.POINTER(Tallies), C.c_int]
        L.orc_run_lanes.argtypes = [C.POINTER(Sim), U64P, U32P, C.c_uint32, C.c_uint64, C.POINTER(Tallies), C.c_int]
        L.orc_run_fluor_jobs.argtypes = [C.POINTER(Sim), U32P, C.c_uint32, C.c_uint32, U32P, U64P, F64P,
                                         C.c_uint64, C.POINTER(Tallies), C.c_int]
        L.orc_run_fluor_lanes.argtypes = [C.POINTER(Sim), U64P, U32P, C.c_uint32, C.c_uint32, U32P, F64P,
                                          C.c_uint64, C.POINTER(Tallies), C.c_int]
        L.orc_fluor_reference.restype = C.c_uint64
        L.orc_fluor_reference.argtypes = [C.POINTER(Sim), F64P, U64P, U32P, C.c_uint32, C.c_uint64, F64P, F64P,
                                          C.c_int]
        L.orc_reflect_test.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, F32P, F32P]
        L.orc_fresnel_test.restype = C.c_float
        L.orc_fresnel_test.argtypes = [C.c_float, C.c_float, C.c_float]
        L.orc_spin_test.argtypes = [C.c_float, F32P, U64P, C.c_uint32, F32P]
        L.orc_backstep_test.argtypes = [F32P, F32P, F32P, C.c_float, F32P, F32P]
        _lib = L
    return _lib


def ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


class Medium:
    """Holds a problem description (arrays kept alive) and builds orc_sim."""

    def __init__(self, labels, bulks, esp, gs, halfx=None, halfy=None):
        self.labels = np.ascontiguousarray(labels, dtype=np.uint8)  # (nz, ny, nx)
        self.nz, self.ny, self.nx = self.labels.shape
        self.bulks_np = np.asarray(bulks, dtype=np.float32).reshape(-1, 8)
        self.bulks = (Bulk * len(self.bulks_np))(*[Bulk(*map(float, r)) for r in self.bulks_np])
        self.esp = np.float32(esp)
        self.gs = float(gs)
        self.halfx = np.float32(2.0) * self.esp if halfx is None else np.float32(halfx)
        self.halfy = np.float32(2.0) * self.esp if halfy is None else np.float32(halfy)
        self.n_bulks = len(self.bulks_np) - 2

    def sim(self, det=(0.1, 0.1, 0.0, 0.0, 10, 10), src=(1, 0.0, 0.0, 0.0, 0.0), start_weight=0xFFFFFFFF,
            fhd=False, axyz=False, sq=False, temporal=None):
        s = Sim()
        s.nx, s.ny, s.nz = self.nx, self.ny, self.nz
        s.esp, s.gs, s.halfx, s.halfy = float(self.esp), self.gs, float(self.halfx), float(self.halfy)
        s.n_bulks = self.n_bulks
        s.bulks = C.cast(self.bulks, C.POINTER(Bulk))
        s.labels = self.labels.ctypes.data_as(C.POINTER(C.c_uint8))
        s.det_dx, s.det_dy, s.det_x0, s.det_y0, s.det_nx, s.det_ny = det
        s.src_kind, s.xi, s.yi, s.zi, s.src_r = src
        s.start_weight = start_weight
        s.fhd_on, s.axyz_on, s.sq_on = int(fhd), int(axyz), int(sq)
        if temporal is not None:
            posx, posy, r, rort, bins, tmax = temporal
            self._tpx = np.asarray(posx, np.float32)
            self._tpy = np.asarray(posy, np.float32)
            s.do_temp, s.t_ndx, s.t_ndy = 1, len(self._tpx), len(self._tpy)
            s.t_posx, s.t_posy = ptr(self._tpx, F32P), ptr(self._tpy, F32P)
            s.t_r, s.t_rort, s.t_bins, s.t_max = r, rort, bins, tmax
        return s


class Result:
    def __init__(self, sim, sq=False):
        npix = sim.det_nx * (sim.det_ny + 1)
        nvox = sim.nx * sim.ny * sim.nz
        self.Rd = np.zeros(npix, np.uint64)
        self.Tt = np.zeros(npix, np.uint64)
        self.fhd = np.zeros(nvox, np.uint64) if sim.fhd_on else None
        self.axyz = np.zeros(nvox, np.uint64) if sim.axyz_on else None
        self.time = np.zeros(max(1, sim.t_ndx * sim.t_ndy * sim.t_bins), np.uint64) if sim.do_temp else None
        self.Rd2 = np.zeros(npix, np.uint64) if sim.sq_on else None
        self.Tt2 = np.zeros(npix, np.uint64) if sim.sq_on else None
        self.t = Tallies()
        self.t.Rd, self.t.Tt = ptr(self.Rd, U64P), ptr(self.Tt, U64P)
        self.t.fhd, self.t.axyz, self.t.time = ptr(self.fhd, U64P), ptr(self.axyz, U64P), ptr(self.time, U64P)
        self.t.Rd2, self.t.Tt2 = ptr(self.Rd2, U64P), ptr(self.Tt2, U64P)

    @property
    def tot(self):
        return {n: int(self.t.tot[i]) for i, n in enumerate(TOT_NAMES)}


def run_photons(sim, a_tab, key, k0, k1, exit_scale=0.0, nthreads=0):
    r = Result(sim)
    a_tab = np.ascontiguousarray(a_tab, np.uint32)
    rc = lib().orc_run_photons(C.byref(sim), ptr(a_tab, U32P), len(a_tab), key, k0, k1, exit_scale,
                               C.byref(r.t), nthreads)
    assert rc == 0, rc
    return r


def run_lanes(sim, x, a, N, nthreads=0):
    r = Result(sim)
    rc = lib().orc_run_lanes(C.byref(sim), ptr(x, U64P), ptr(a, U32P), len(a), N, C.byref(r.t), nthreads)
    if rc != 0:
        raise ValueError(f"orc_run_lanes rc={rc}")
    return r


def init_rng(n, primes, seed):
    x = np.zeros(n, np.uint64)
    a = np.zeros(n, np.uint32)
    rc = lib().orc_init_rng(ptr(x, U64P), ptr(a, U32P), n, primes.encode(), seed)
    return rc, x, a


def read_primes(path, n=None):
    vals = np.loadtxt(path, dtype=np.uint64, usecols=0, max_rows=None if n is None else n + 1)
    return vals.astype(np.uint32)


def tdet_positions(n, sep, x0):
    """Temporal detector positions in float32, same operation order as the
    reference's DoOneSimulation (CUDAMCFLmain.cu:95-100)."""
    f = np.float32
    return np.array([f(f(f(i) * f(sep)) - f(f(f(n - 1) * f(sep)) / f(2))) + f(x0) for i in range(n)], np.float32)


class FromProduct:
    """orc_sim built from a cudamcfl_amd.Simulation (same arrays, same numbers)."""

    def __init__(self, sim, emission=False, axyz=False, sq=False):
        s = sim.s
        o = Sim()
        if s.bulk_method == 1:
            # layered slab: layer optics as bulks, their bounds, the inclusion
            nb = s.n_layers + 2
            Ls = [s.layers[i] for i in range(nb)]
            rows = [[L.mutr, L.mua, L.g, L.n, L.flc, L.muaf, L.eY, L.albedof] for L in Ls]
            self.zmin = np.array([L.z_min for L in Ls], np.float32)
            self.zmax = np.array([L.z_max for L in Ls], np.float32)
            c = s.inclusion
            o.layered = 1
            o.z_min, o.z_max = ptr(self.zmin, F32P), ptr(self.zmax, F32P)
            o.inc = Bulk(c.mutr, c.mua, c.g, c.n, c.flc, c.muaf, c.eY, c.albedof)
            o.inc_x, o.inc_y, o.inc_z, o.inc_r = c.x, c.y, c.z, c.r
            o.start_layer = next((l for l in range(1, s.n_layers + 1)
                                  if self.zmin[l] <= np.float32(s.zi) < self.zmax[l]), 0)
            self.labels = np.zeros(1, np.uint8)
            o.n_bulks = s.n_layers
        else:
            nb = s.n_bulks + 2
            rows = []
            for i in range(nb):
                b = s.bulks[i]
                mua = b.muaf if emission else b.mua
                rows.append([b.mutr, mua, b.g, b.n, b.flc, b.muaf, b.eY, b.albedof])
            self.labels = np.array(sim.labels, copy=True)
            o.n_bulks = s.n_bulks
        self.bulks = (Bulk * nb)(*[Bulk(*map(float, r)) for r in np.asarray(rows, np.float32)])
        o.nx, o.ny, o.nz = s.nx, s.ny, s.nz
        o.esp, o.gs, o.halfx, o.halfy = s.esp, float(s.grid_size), s.halfx, s.halfy
        o.bulks = C.cast(self.bulks, C.POINTER(Bulk))
        o.labels = self.labels.ctypes.data_as(C.POINTER(C.c_uint8))
        d = s.det
        o.det_dx, o.det_dy, o.det_x0, o.det_y0, o.det_nx, o.det_ny = d.dx, d.dy, d.x0, d.y0, d.nx, d.ny
        if s.do_temp_sim == 1 and not emission:
            self._tpx = tdet_positions(d.x_temp_numdets, d.x_temp_sepdets, d.x0_temp_det)
            self._tpy = tdet_positions(d.y_temp_numdets, d.y_temp_sepdets, d.y0_temp_det)
            o.do_temp, o.t_ndx, o.t_ndy = 1, d.x_temp_numdets, d.y_temp_numdets
            o.t_posx, o.t_posy = ptr(self._tpx, F32P), ptr(self._tpy, F32P)
            o.t_r, o.t_rort, o.t_bins, o.t_max = d.temp_det_r, d.temp_rort, d.temp_bins, d.max_temp
        o.src_kind = 0 if s.dir == 0.0 else (2 if s.src_kind == 2 else 1)
        o.xi, o.yi, o.zi, o.src_r = s.xi, s.yi, s.zi, s.src_r
        o.start_weight = s.start_weight
        o.fhd_on = 1 if (s.fhd_activated == 1 and not emission) else 0
        o.axyz_on, o.sq_on = int(axyz), int(sq)
        self.o = o


def run_fluor_jobs(osim, a_tab, vox, keys, scale, n_per_job, nthreads=0):
    r = Result(osim)
    vox = np.ascontiguousarray(vox, np.uint32)
    keys = np.ascontiguousarray(keys, np.uint64)
    scale = np.ascontiguousarray(scale, np.float64)
    a_tab = np.ascontiguousarray(a_tab, np.uint32)
    rc = lib().orc_run_fluor_jobs(C.byref(osim), ptr(a_tab, U32P), len(a_tab), len(vox), ptr(vox, U32P),
                                  ptr(keys, U64P), ptr(scale, F64P), n_per_job, C.byref(r.t), nthreads)
    assert rc == 0, rc
    return r


def fluor_reference(osim, Fx, x, a, n_per_voxel, nthreads=0):
    """The reference's emission loop (every emitting voxel, n_per_voxel
    photons on the continuing lane streams x, a; x advanced in place).
    Returns (FlR, FlT, photons simulated)."""
    xy = osim.det_nx * osim.det_ny
    FlR, FlT = np.zeros(xy), np.zeros(xy)
    Fx = np.ascontiguousarray(Fx, np.float64)
    n = lib().orc_fluor_reference(C.byref(osim), ptr(Fx, F64P), ptr(x, U64P), ptr(a, U32P), len(a), n_per_voxel,
                                  ptr(FlR, F64P), ptr(FlT, F64P), nthreads)
    return FlR, FlT, int(n)


def run_fluor_lanes(osim, x, a, vox, scale, n_per_voxel, nthreads=0):
    """The reference's emission loop on its lanes, fixed-point images (x advanced in place)."""
    r = Result(osim)
    vox = np.ascontiguousarray(vox, np.uint32)
    scale = np.ascontiguousarray(scale, np.float64)
    rc = lib().orc_run_fluor_lanes(C.byref(osim), ptr(x, U64P), ptr(a, U32P), len(a), len(vox), ptr(vox, U32P),
                                   ptr(scale, F64P), n_per_voxel, C.byref(r.t), nthreads)
    if rc == -2:
        raise ValueError("n_per_voxel < lanes")
    assert rc == 0, rc
    return r


def _bind_extras(L):
    if getattr(L, "_extras", False):
        return
    L.orc_math_eval.argtypes = [C.c_int, F32P, F32P, C.c_size_t]
    L.orc_mwc_draws.argtypes = [C.c_uint64, C.c_uint32, C.c_size_t, U64P, F32P]
    L.orc_photon_state.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32]
    L.orc_photon_state.restype = C.c_uint64
    L.orc_tof_div_mismatches.argtypes = [C.c_uint32, C.c_uint32]
    L.orc_tof_div_mismatches.restype = C.c_uint64
    L.orc_set_max_steps.argtypes = [C.c_uint32]
    L.orc_launch_sample.argtypes = [C.POINTER(Sim), C.c_int, U32P, C.c_uint32, C.c_uint64, C.c_uint64, F32P, U32P,
                                    U32P]
    L.orc_launch_sample.restype = None
    L.orc_survive_test.argtypes = [C.c_uint32, U64P, C.c_uint32, U32P, U64P, U64P]
    L.orc_march_test.argtypes = [C.POINTER(Sim), F32P, C.c_uint32, C.c_float, F32P, U64P]
    L.orc_march_test.restype = C.c_uint32
    L.orc_survive_batch.argtypes = [C.c_uint32, U32P, C.c_uint32, C.c_uint64, C.c_uint64, U64P, U64P]
    L.orc_survive_batch.restype = None
    L._extras = True


def survive_batch(w, a_tab, key, n):
    """PhotonSurvive on n photon streams of weight w: (photons that live, sum of their weights)."""
    L = lib()
    _bind_extras(L)
    a_tab = np.ascontiguousarray(a_tab, np.uint32)
    lived = np.zeros(1, np.uint64)
    sw = np.zeros(1, np.uint64)
    L.orc_survive_batch(w, ptr(a_tab, U32P), len(a_tab), key, n, ptr(lived, U64P), ptr(sw, U64P))
    return int(lived[0]), int(sw[0])


def launch_sample(sim, kind, a_tab, key, n):
    """n photons of the oracle's LaunchPhoton on photon streams: (pos_dir
    (n, 6) float32, weight, entry label)."""
    L = lib()
    _bind_extras(L)
    a_tab = np.ascontiguousarray(a_tab, np.uint32)
    pd = np.zeros((n, 6), np.float32)
    w = np.zeros(n, np.uint32)
    lab = np.zeros(n, np.uint32)
    L.orc_launch_sample(C.byref(sim), kind, ptr(a_tab, U32P), len(a_tab), key, n, ptr(pd, F32P), ptr(w, U32P),
                        ptr(lab, U32P))
    return pd, w, lab


def survive(w, x, a):
    """The oracle's PhotonSurvive: (lives, weight after, state after, gain, loss)."""
    L = lib()
    _bind_extras(L)
    xs = np.array([x], np.uint64)
    wo = np.zeros(1, np.uint32)
    g = np.zeros(1, np.uint64)
    lo = np.zeros(1, np.uint64)
    r = L.orc_survive_test(w, ptr(xs, U64P), a, ptr(wo, U32P), ptr(g, U64P), ptr(lo, U64P))
    return r, int(wo[0]), int(xs[0]), int(g[0]), int(lo[0])


def march(sim, pos, direction, old_label, max_s=100.0):
    """The oracle's MoveToFirstBoundary: (label found, position after, time of flight added)."""
    L = lib()
    _bind_extras(L)
    pd = np.array(list(pos) + list(direction), np.float32)
    out = np.zeros(3, np.float32)
    tof = np.zeros(1, np.uint64)
    lab = L.orc_march_test(C.byref(sim), ptr(pd, F32P), old_label, max_s, ptr(out, F32P), ptr(tof, U64P))
    return int(lab), out, int(tof[0])


def math_eval(fn, x):
    L = lib()
    _bind_extras(L)
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros_like(x)
    L.orc_math_eval(fn, ptr(x, F32P), ptr(out, F32P), len(x))
    return out


def mwc_draws(x, a, n):
    L = lib()
    _bind_extras(L)
    st = np.zeros(n, np.uint64)
    co = np.zeros(n, np.float32)
    L.orc_mwc_draws(x, a, n, ptr(st, U64P), ptr(co, F32P))
    return st, co


def photon_state(key, k, a):
    L = lib()
    _bind_extras(L)
    return int(L.orc_photon_state(key, k, a))


def set_max_steps(n):
    """The oracle's per-photon step cap for later runs (0: the default)."""
    L = lib()
    _bind_extras(L)
    L.orc_set_max_steps(int(n))


def tof_div_mismatches(lo, hi):
    L = lib()
    _bind_extras(L)
    return int(L.orc_tof_div_mismatches(lo, hi))


def reference_order(sim, vox, scale):
    """Emission voxels (x fastest, as fl_plan lists them) in the reference
    loop's order: x outermost, then y, then z (CUDAMCFLmain.cu:471-473)."""
    nx, ny = sim.s.nx, sim.s.ny
    v = np.asarray(vox, np.int64)
    ix, iy, iz = v % nx, (v // nx) % ny, v // (nx * ny)
    order = np.lexsort((iz, iy, ix))
    return np.asarray(vox)[order], np.asarray(scale)[order]
