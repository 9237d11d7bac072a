"""cudamcfl_amd -- MI355X voxelized fluorescence Monte Carlo (Python binding).

Thin ctypes mirror of the C ABI in include/mcfl.h (libmcfl.so, built in-tree
into cudamcfl_amd/_build/).  The structures keep the reference's field names
(SimulationStruct, BulkStruct, DetStruct of CUDAMCFL.h) so code written
against the reference's driver reads the same here.  There is no CPU fallback:
if the library is missing, loading fails loudly.
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
BUILD_DIR = os.path.join(PKG_DIR, "_build")
LIB_PATH = os.path.join(BUILD_DIR, "libmcfl.so")
PRIMES_PATH = os.path.join(PKG_DIR, "data", "safeprimes_base32.txt")

STR_LEN = 500
NUM_THREADS = 32768
NTOT = 16
TOT_NAMES = ["launched_w", "R", "T", "A", "roul_gain", "roul_loss", "zero_kill", "photons", "steps", "aborted"]

SOURCE_ISOTROPIC, SOURCE_PENCIL, SOURCE_DISK = 0, 1, 2
RNG_PHOTON, RNG_LANES = 0, 1

U64P = C.POINTER(C.c_uint64)
U32P = C.POINTER(C.c_uint32)
F64P = C.POINTER(C.c_double)


class Layer(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("z_min", "z_max", "mutr", "mua", "g", "n", "flc", "muaf", "eY", "albedof")]


class Bulk(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("mutr", "mua", "g", "n", "flc", "muaf", "eY", "albedof")]


class Inclusion(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("x", "y", "z", "r", "mutr", "mua", "g", "n", "flc", "muaf", "eY",
                                         "albedof")]


class Det(C.Structure):
    _fields_ = [("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
                ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("x0", C.c_float), ("y0", C.c_float), ("sep", C.c_float),
                ("x0_temp_det", C.c_float), ("y0_temp_det", C.c_float),
                ("x_temp_numdets", C.c_uint32), ("y_temp_numdets", C.c_uint32),
                ("x_temp_sepdets", C.c_float), ("y_temp_sepdets", C.c_float), ("temp_det_r", C.c_float),
                ("temp_rort", C.c_uint32), ("temp_bins", C.c_uint64), ("max_temp", C.c_uint64)]


class Sim(C.Structure):
    _fields_ = [("number_of_photons", C.c_uint64), ("number_of_photons_per_voxel", C.c_uint32),
                ("n_layers", C.c_uint32), ("n_bulks", C.c_uint32), ("start_weight", C.c_uint32),
                ("outp_filename", C.c_char * STR_LEN), ("inp_filename", C.c_char * STR_LEN),
                ("begin", C.c_int64), ("end", C.c_int64), ("AorB", C.c_char),
                ("det", Det), ("layers", C.POINTER(Layer)), ("bulks", C.POINTER(Bulk)), ("inclusion", Inclusion),
                ("grid_size", C.c_int32), ("esp", C.c_float), ("xi", C.c_float), ("yi", C.c_float),
                ("zi", C.c_float), ("dir", C.c_float), ("n_down", C.c_float),
                ("fhd_activated", C.c_int32), ("do_fl_sim", C.c_int32), ("bulk_method", C.c_int32),
                ("bulk_info", C.POINTER(C.c_uint8)), ("bulkinfo_filename", C.c_char * STR_LEN),
                ("do_temp_sim", C.c_int32),
                ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("halfx", C.c_float), ("halfy", C.c_float),
                ("src_kind", C.c_int32), ("src_r", C.c_float)]


class Rng(C.Structure):
    _fields_ = [("mode", C.c_int32), ("seed", C.c_uint64), ("x", U64P), ("a", U32P), ("n_lanes", C.c_uint32),
                ("first_photon", C.c_uint64)]


class Tallies(C.Structure):
    _fields_ = [("Rd_xy", U64P), ("Tt_xy", U64P), ("fhd", U64P), ("time_xyt", U64P), ("A_xyz", U64P),
                ("Rd2_xy", U64P), ("Tt2_xy", U64P), ("totals", C.c_uint64 * NTOT),
                ("num_terminated_photons", C.c_uint64), ("kernel_ms", C.c_double)]


class FlResult(C.Structure):
    _fields_ = [("FlR_raw", U64P), ("FlT_raw", U64P), ("unit", C.c_double), ("voxels", C.c_uint64),
                ("photons", C.c_uint64), ("totals", C.c_uint64 * NTOT), ("kernel_ms", C.c_double)]


class McflError(RuntimeError):
    pass


_lib = None

EXPORTS = ["mcfl_read_simulation_data", "mcfl_free_simulations", "mcfl_interpret_arg", "mcfl_init_rng",
           "mcfl_default_primes_path", "mcfl_stream_key", "mcfl_ctx_create", "mcfl_ctx_destroy",
           "mcfl_status_string", "mcfl_ctx_info", "mcfl_run_excitation", "mcfl_DoOneSimulation",
           "mcfl_DoOneSimulationFl", "mcfl_run_excitation_noalloc", "mcfl_fl_plan",
           "mcfl_run_fluorescence", "mcfl_run_fluorescence_lanes", "mcfl_write_simulation_results", "mcfl_write_phd",
           "mcfl_write_fl"]


def lib():
    """Load libmcfl.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise McflError(f"libmcfl.so not built at {LIB_PATH}; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    L.mcfl_read_simulation_data.argtypes = [C.c_char_p, C.POINTER(C.POINTER(Sim)), C.c_int]
    L.mcfl_free_simulations.argtypes = [C.POINTER(Sim), C.c_int]
    L.mcfl_free_simulations.restype = None
    L.mcfl_interpret_arg.argtypes = [C.c_int, C.POINTER(C.c_char_p), U64P, C.POINTER(C.c_int)]
    L.mcfl_init_rng.argtypes = [U64P, U32P, C.c_uint32, C.c_char_p, C.c_uint64]
    L.mcfl_default_primes_path.restype = C.c_char_p
    L.mcfl_stream_key.argtypes = [C.c_uint64, C.c_uint64]
    L.mcfl_stream_key.restype = C.c_uint64
    L.mcfl_ctx_create.argtypes = [C.c_int, C.c_char_p, C.POINTER(C.c_void_p)]
    L.mcfl_ctx_destroy.argtypes = [C.c_void_p]
    L.mcfl_ctx_destroy.restype = None
    L.mcfl_status_string.argtypes = [C.c_int]
    L.mcfl_status_string.restype = C.c_char_p
    L.mcfl_ctx_info.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_char_p, C.c_int]
    L.mcfl_run_excitation.argtypes = [C.c_void_p, C.POINTER(Sim), C.POINTER(Rng), C.POINTER(Tallies)]
    L.mcfl_DoOneSimulation.argtypes = [C.c_void_p, C.POINTER(Sim), U64P, U32P, F64P]
    L.mcfl_DoOneSimulation.restype = C.c_uint64
    L.mcfl_run_excitation_noalloc.argtypes = [C.c_void_p, C.POINTER(Sim), C.POINTER(Rng), C.POINTER(Tallies)]
    L.mcfl_DoOneSimulationFl.argtypes = [C.c_void_p, C.POINTER(Sim), U64P, U32P, U64P, U64P]
    L.mcfl_DoOneSimulationFl.restype = C.c_uint64
    L.mcfl_fl_plan.argtypes = [C.POINTER(Sim), F64P, C.c_uint64, U32P, F64P, U64P, F64P]
    L.mcfl_run_fluorescence.argtypes = [C.c_void_p, C.POINTER(Sim), F64P, C.c_uint64, C.c_uint64, F64P, F64P,
                                        C.POINTER(FlResult)]
    L.mcfl_run_fluorescence_lanes.argtypes = [C.c_void_p, C.POINTER(Sim), F64P, C.c_uint64, C.POINTER(Rng), F64P, F64P,
                                              C.POINTER(FlResult)]
    L.mcfl_write_simulation_results.argtypes = [C.POINTER(Sim), C.POINTER(Tallies), C.c_double]
    L.mcfl_write_phd.argtypes = [C.POINTER(Sim), F64P]
    L.mcfl_write_fl.argtypes = [C.POINTER(Sim), F64P, F64P]
    _lib = L
    return L


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def status_string(st):
    return lib().mcfl_status_string(st).decode()


def check(st, what):
    if st != 0:
        raise McflError(f"{what}: {status_string(st)} ({st})")


def stream_key(seed, stream):
    return int(lib().mcfl_stream_key(seed, stream))


def init_rng(n=NUM_THREADS, seed=1, primes=None):
    """init_RNG (CUDAMCFLrng.cu:31-90): per-lane MWC states x and multipliers a."""
    x = np.zeros(n, np.uint64)
    a = np.zeros(n, np.uint32)
    rc = lib().mcfl_init_rng(_p(x, U64P), _p(a, U32P), n, (primes or "").encode(), seed)
    if rc != 0:
        raise McflError(f"init_rng failed (seed {seed})")
    return x, a


def multipliers(n=32768, primes=None):
    """Per-lane multipliers (column 1 of lines 2..n+1 of the primes file)."""
    vals = np.loadtxt(primes or PRIMES_PATH, dtype=np.uint64, usecols=0, max_rows=n + 1)
    return vals[1:].astype(np.uint32)


class Simulation:
    """Owns an mcfl_sim plus the arrays it points to."""

    def __init__(self):
        self.s = Sim()
        self._keep = []
        self._owned_by_lib = None

    @classmethod
    def from_mci(cls, path, index=0):
        """read_simulation_data (CUDAMCFLio.cu:243-659)."""
        sims = C.POINTER(Sim)()
        n = lib().mcfl_read_simulation_data(os.fsencode(path), C.byref(sims), 0)
        if n == 0:
            raise McflError(f"failed to parse {path}")
        obj = cls()
        obj.s = sims[index]
        obj._owned_by_lib = (sims, n)
        obj.n_runs = n
        return obj

    @classmethod
    def voxel(cls, labels, bulks, esp, grid_size, det=(0.1, 0.1, 10, 10, 0.0, 0.0), photons=100000,
              source=(0.0, 0.0, 0.0), src_kind=SOURCE_PENCIL, src_r=0.0, n_above=1.0, n_below=1.0,
              fhd=False, reference_grid=False, temporal=None, photons_per_voxel=40000, outp="mcfl-out"):
        """A bulk-method-2 simulation from arrays.

        labels: uint8 (nz, ny, nx); bulks: rows (n, mua, mus, flc, AbsCExi, AbsCEmi, eY, g) as in the
        .mci medium lines (CUDAMCFLio.cu:557-568), derived optics computed like the parser.
        """
        obj = cls()
        s = obj.s
        lab = np.ascontiguousarray(labels, dtype=np.uint8)
        nz, ny, nx = lab.shape
        rows = np.asarray(bulks, dtype=np.float32).reshape(-1, 8)
        nb = len(rows)
        B = (Bulk * (nb + 2))()
        f32 = np.float32

        def outer(b, n):
            b.n, b.mua, b.muaf, b.eY, b.flc, b.g, b.mutr, b.albedof = n, 0, 0, 0, 0, 1, np.finfo(f32).max, 0

        outer(B[0], n_above)
        outer(B[nb + 1], n_below)
        for i, r in enumerate(rows):
            n, mua, mus, flc, exi, emi, ey, g = [f32(v) for v in r]
            b = B[i + 1]
            b.n = n
            b.mua = f32(mua + f32(flc * exi))
            b.muaf = f32(mua + f32(flc * emi))
            b.eY, b.flc, b.g = ey, flc, g
            b.mutr = np.finfo(f32).max if mus == 0 else f32(f32(1) / f32(mua + mus))
            with np.errstate(invalid="ignore"):  # 0/0 = NaN for a glass without fluorophore, as the parser gives
                b.albedof = f32(mus / f32(f32(flc * exi) + mus))
        obj._keep += [lab, B]
        s.number_of_photons = photons
        s.number_of_photons_per_voxel = photons_per_voxel
        s.n_bulks = nb
        s.bulks = C.cast(B, C.POINTER(Bulk))
        s.bulk_info = lab.ctypes.data_as(C.POINTER(C.c_uint8))
        s.bulk_method = 2
        s.grid_size = int(grid_size)
        s.esp = f32(esp)
        s.nx, s.ny, s.nz = nx, ny, nz
        if reference_grid:
            s.halfx = s.halfy = f32(2) * f32(esp)
        else:
            s.halfx = f32(nx) / (f32(2) * f32(grid_size))
            s.halfy = f32(ny) / (f32(2) * f32(grid_size))
        s.det.dx, s.det.dy, s.det.nx, s.det.ny, s.det.x0, s.det.y0 = det
        s.xi, s.yi, s.zi = source
        s.dir = 1.0 if src_kind != SOURCE_ISOTROPIC else 0.0
        s.src_kind, s.src_r = src_kind, src_r
        s.fhd_activated = 1 if fhd else 0
        s.do_fl_sim = 1 if fhd else 0
        r = (float(n_above) - float(rows[0][0])) / (float(n_above) + float(rows[0][0]))
        s.start_weight = int(float(0xFFFFFFFF) * (1 - r * r))
        s.outp_filename = outp.encode()
        if temporal is not None:
            (x0, y0, ndx, ndy, sx, sy, rad, rort, bins, tmax) = temporal
            s.do_temp_sim = 1
            d = s.det
            d.x0_temp_det, d.y0_temp_det, d.x_temp_numdets, d.y_temp_numdets = x0, y0, ndx, ndy
            d.x_temp_sepdets, d.y_temp_sepdets, d.temp_det_r, d.temp_rort = sx, sy, rad, rort
            d.temp_bins, d.max_temp = bins, tmax
        return obj

    @classmethod
    def layered(cls, layers, inclusion, grid_size, det=(0.1, 0.1, 10, 10, 0.0, 0.0), photons=100000,
                source=(0.0, 0.0, 0.0), src_kind=SOURCE_PENCIL, src_r=0.0, n_above=1.0, n_below=1.0, fhd=False,
                temporal=None, outp="mcfl-out"):
        """A bulk-method-1 simulation: a layered slab with one spherical inclusion.

        layers: rows (n, mua, mus, flc, AbsCExi, AbsCEmi, eY, g, d) as in the .mci layer lines;
        inclusion: (x, y, z, r, n, mua, mus, flc, AbsCExi, AbsCEmi, eY, g); derived optics,
        layer bounds and the fhd grid computed like the parser (CUDAMCFLio.cu:427-526).
        """
        obj = cls()
        s = obj.s
        f32 = np.float32
        rows = np.asarray(layers, dtype=np.float32).reshape(-1, 9)
        nl = len(rows)
        Ls = (Layer * (nl + 2))()
        Ls[0].n = f32(n_above)
        Ls[nl + 1].n = f32(n_below)
        dtot = f32(0)
        for i, r in enumerate(rows):
            n, mua, mus, flc, exi, emi, ey, g, d = [f32(v) for v in r]
            L = Ls[i + 1]
            L.n = n
            L.mua = f32(mua + f32(flc * exi))
            L.muaf = f32(mua + f32(flc * emi))
            L.eY, L.flc, L.g = ey, flc, g
            L.z_min = dtot
            dtot = f32(dtot + d)
            L.z_max = dtot
            L.mutr = np.finfo(f32).max if mus == 0 else f32(f32(1) / f32(mua + mus))
            with np.errstate(invalid="ignore"):  # 0/0 = NaN, as the parser gives
                L.albedof = f32(mus / f32(f32(flc * exi) + mus))
        x, y, z, r, n, mua, mus, flc, exi, emi, ey, g = [f32(v) for v in inclusion]
        inc = s.inclusion
        inc.x, inc.y, inc.z, inc.r, inc.n = x, y, z, r, n
        inc.mua = f32(mua + f32(flc * exi))
        inc.muaf = f32(mua + f32(flc * emi))
        inc.flc, inc.eY, inc.g = flc, ey, g
        inc.mutr = np.finfo(f32).max if mus == 0 else f32(f32(1) / f32(mua + mus))
        with np.errstate(invalid="ignore", divide="ignore"):  # 0/0 as the parser computes it
            inc.albedof = f32(mus / f32(f32(flc * exi) + mus))
        B = (Bulk * 1)()
        lab = np.zeros(1, np.uint8)
        obj._keep += [Ls, B, lab]
        s.number_of_photons = photons
        s.n_layers = nl
        s.layers = C.cast(Ls, C.POINTER(Layer))
        s.n_bulks = 0
        s.bulks = C.cast(B, C.POINTER(Bulk))
        s.bulk_info = lab.ctypes.data_as(C.POINTER(C.c_uint8))
        s.bulk_method = 1
        s.grid_size = int(grid_size)
        s.esp = dtot
        s.nx = s.ny = int(f32(f32(4) * dtot) * f32(grid_size))
        s.nz = int(f32(dtot * f32(grid_size)))
        s.halfx = s.halfy = f32(2) * dtot
        s.det.dx, s.det.dy, s.det.nx, s.det.ny, s.det.x0, s.det.y0 = det
        s.xi, s.yi, s.zi = source
        s.dir = 1.0 if src_kind != SOURCE_ISOTROPIC else 0.0
        s.src_kind, s.src_r = src_kind, src_r
        s.fhd_activated = 1 if fhd else 0
        rr = (float(n_above) - float(rows[0][0])) / (float(n_above) + float(rows[0][0]))
        s.start_weight = int(float(0xFFFFFFFF) * (1 - rr * rr))
        s.outp_filename = outp.encode()
        if temporal is not None:
            (x0, y0, ndx, ndy, sx, sy, rad, rort, bins, tmax) = temporal
            s.do_temp_sim = 1
            d = s.det
            d.x0_temp_det, d.y0_temp_det, d.x_temp_numdets, d.y_temp_numdets = x0, y0, ndx, ndy
            d.x_temp_sepdets, d.y_temp_sepdets, d.temp_det_r, d.temp_rort = sx, sy, rad, rort
            d.temp_bins, d.max_temp = bins, tmax
        return obj

    @property
    def nvox(self):
        return self.s.nx * self.s.ny * self.s.nz

    @property
    def npix(self):
        return self.s.det.nx * (self.s.det.ny + 1)

    @property
    def labels(self):
        return np.ctypeslib.as_array(self.s.bulk_info, shape=(self.nvox,))

    def __del__(self):
        if self._owned_by_lib is not None and _lib is not None:
            sims, n = self._owned_by_lib
            self._owned_by_lib = None
            _lib.mcfl_free_simulations(sims, n)


class ExcitationResult:
    def __init__(self, sim, fhd=True, axyz=False, sq=False, temporal=True):
        s = sim.s
        self.Rd = np.zeros(sim.npix, np.uint64)
        self.Tt = np.zeros(sim.npix, np.uint64)
        self.fhd = np.zeros(sim.nvox, np.uint64) if (fhd and s.fhd_activated == 1) else None
        self.A_xyz = np.zeros(sim.nvox, np.uint64) if axyz else None
        nt = s.det.x_temp_numdets * s.det.y_temp_numdets * s.det.temp_bins
        self.time = np.zeros(max(nt, 1), np.uint64) if (temporal and s.do_temp_sim == 1) else None
        self.Rd2 = np.zeros(sim.npix, np.uint64) if sq else None
        self.Tt2 = np.zeros(sim.npix, np.uint64) if sq else None
        t = Tallies()
        t.Rd_xy, t.Tt_xy = _p(self.Rd, U64P), _p(self.Tt, U64P)
        t.fhd, t.A_xyz, t.time_xyt = _p(self.fhd, U64P), _p(self.A_xyz, U64P), _p(self.time, U64P)
        t.Rd2_xy, t.Tt2_xy = _p(self.Rd2, U64P), _p(self.Tt2, U64P)
        self.t = t

    @property
    def totals(self):
        return {n: int(self.t.totals[i]) for i, n in enumerate(TOT_NAMES)}

    @property
    def kernel_ms(self):
        return self.t.kernel_ms


class Context:
    """A device context (mcfl_ctx): one GPU, its stream and resident buffers."""

    def __init__(self, device=0, primes=None):
        import weakref
        self.h = C.c_void_p()
        self._children = weakref.WeakSet()  # batches and servers: destroyed before the context
        check(lib().mcfl_ctx_create(device, (primes or "").encode(), C.byref(self.h)), "mcfl_ctx_create")

    def close(self):
        if self.h:
            for ch in list(self._children):
                ch.close()
            lib().mcfl_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        ncu, bpc = C.c_int(), C.c_int()
        name = C.create_string_buffer(256)
        check(lib().mcfl_ctx_info(self.h, C.byref(ncu), C.byref(bpc), name, 256), "mcfl_ctx_info")
        return {"n_cu": ncu.value, "blocks_per_cu": bpc.value, "arch": name.value.decode()}

    def run_excitation(self, sim, seed=None, lanes=None, axyz=False, sq=False, first_photon=0, noalloc=False):
        """mcfl_run_excitation.  seed -> photon streams (photons first_photon onwards);
        lanes=(x, a) -> reference lanes (x advanced).  noalloc: the
        allocation-free variant mcfl_run_excitation_noalloc."""
        r = ExcitationResult(sim, axyz=axyz, sq=sq)
        rng = Rng()
        if lanes is not None:
            x, a = lanes
            rng.mode, rng.x, rng.a, rng.n_lanes = RNG_LANES, _p(x, U64P), _p(a, U32P), len(a)
        else:
            rng.mode, rng.seed = RNG_PHOTON, int(seed if seed is not None else 1)
        rng.first_photon = int(first_photon)  # the library refuses an offset on lane streams
        fn = lib().mcfl_run_excitation_noalloc if noalloc else lib().mcfl_run_excitation
        check(fn(self.h, C.byref(sim.s), C.byref(rng), C.byref(r.t)), "run_excitation")
        return r

    def DoOneSimulationFl(self, sim, x, a):
        """mcfl_DoOneSimulationFl: one emission simulation of the reference's
        voxel loop on lanes (x, a) (x advanced in place), with sim as the
        caller set it up.  Returns (photons finished, R, T); 0 photons on
        failure."""
        R = np.zeros(sim.npix, np.uint64)
        T = np.zeros(sim.npix, np.uint64)
        n = lib().mcfl_DoOneSimulationFl(self.h, C.byref(sim.s), _p(x, U64P), _p(a, U32P), _p(R, U64P), _p(T, U64P))
        return int(n), R, T

    def run_fluorescence_lanes(self, sim, Fx, n_per_voxel, lanes):
        """mcfl_run_fluorescence_lanes: the voxel loop on the reference's lane streams
        lanes=(x, a) (x advanced in place)."""
        x, a = lanes
        xy = sim.s.det.nx * sim.s.det.ny
        FlR, FlT = np.zeros(xy), np.zeros(xy)
        res = FlResult()
        rawR, rawT = np.zeros(sim.npix, np.uint64), np.zeros(sim.npix, np.uint64)
        res.FlR_raw, res.FlT_raw = _p(rawR, U64P), _p(rawT, U64P)
        rng = Rng()
        rng.mode, rng.x, rng.a, rng.n_lanes = RNG_LANES, _p(x, U64P), _p(a, U32P), len(a)
        Fx = np.ascontiguousarray(Fx, np.float64)
        check(lib().mcfl_run_fluorescence_lanes(self.h, C.byref(sim.s), _p(Fx, F64P), n_per_voxel, C.byref(rng),
                                                _p(FlR, F64P), _p(FlT, F64P), C.byref(res)), "run_fluorescence_lanes")
        return {"FlR": FlR, "FlT": FlT, "raw_R": rawR, "raw_T": rawT, "unit": res.unit, "voxels": res.voxels,
                "photons": res.photons, "kernel_ms": res.kernel_ms,
                "totals": {n: int(res.totals[i]) for i, n in enumerate(TOT_NAMES)}}

    def run_fluorescence(self, sim, Fx, n_per_voxel, seed):
        xy = sim.s.det.nx * sim.s.det.ny
        FlR = np.zeros(xy)
        FlT = np.zeros(xy)
        res = FlResult()
        rawR = np.zeros(sim.npix, np.uint64)
        rawT = np.zeros(sim.npix, np.uint64)
        res.FlR_raw, res.FlT_raw = _p(rawR, U64P), _p(rawT, U64P)
        Fx = np.ascontiguousarray(Fx, np.float64)
        check(lib().mcfl_run_fluorescence(self.h, C.byref(sim.s), _p(Fx, F64P), n_per_voxel, seed, _p(FlR, F64P),
                                          _p(FlT, F64P), C.byref(res)), "run_fluorescence")
        return {"FlR": FlR, "FlT": FlT, "raw_R": rawR, "raw_T": rawT, "unit": res.unit, "voxels": res.voxels,
                "photons": res.photons, "kernel_ms": res.kernel_ms,
                "totals": {n: int(res.totals[i]) for i, n in enumerate(TOT_NAMES)}}


def fl_plan(sim, Fx, n_per_voxel):
    """mcfl_fl_plan: emitting voxels, per-voxel scale (voxelw/unit) and the unit."""
    cap = C.c_uint64(sim.nvox)
    vox = np.zeros(sim.nvox, np.uint32)
    scale = np.zeros(sim.nvox, np.float64)
    unit = C.c_double()
    Fx = np.ascontiguousarray(Fx, np.float64)
    check(lib().mcfl_fl_plan(C.byref(sim.s), _p(Fx, F64P), n_per_voxel, _p(vox, U32P), _p(scale, F64P),
                             C.byref(cap), C.byref(unit)), "fl_plan")
    n = cap.value
    return vox[:n].copy(), scale[:n].copy(), unit.value


# ------------------------------------------------------------ batch / serving

SOURCE_EMISSION = 3
TALLY_FHD, TALLY_SQ, TALLY_TEMP, TALLY_AXYZ = 1, 2, 4, 8
MAX_BATCH_JOBS = 64


class Job(C.Structure):
    _fields_ = [("medium", C.c_int32), ("src_kind", C.c_int32), ("xi", C.c_float), ("yi", C.c_float),
                ("zi", C.c_float), ("src_r", C.c_float), ("photons", C.c_uint64), ("seed", C.c_uint64),
                ("tallies", C.c_uint32), ("Fx", F64P),
                ("Rd_xy", U64P), ("Tt_xy", U64P), ("fhd", U64P), ("A_xyz", U64P), ("Rd2_xy", U64P),
                ("Tt2_xy", U64P), ("time_xyt", U64P), ("totals", C.c_uint64 * NTOT), ("fl_unit", C.c_double)]


class Request(C.Structure):
    _fields_ = [("medium", C.c_int32), ("src_kind", C.c_int32), ("xi", C.c_float), ("yi", C.c_float),
                ("zi", C.c_float), ("src_r", C.c_float), ("photons", C.c_uint64), ("seed", C.c_uint64),
                ("fluorescence", C.c_int32)]


class RequestResult(C.Structure):
    _fields_ = [("Rd_xy", U64P), ("Tt_xy", U64P), ("FlR_raw", U64P), ("FlT_raw", U64P), ("fl_unit", C.c_double),
                ("totals", C.c_uint64 * NTOT), ("fl_totals", C.c_uint64 * NTOT), ("epochs", C.c_uint64)]


def _bind_engine(L):
    if getattr(L, "_engine", False):
        return
    SP = C.POINTER(C.POINTER(Sim))
    L.mcfl_batch_create.argtypes = [C.c_void_p, SP, C.c_int, C.POINTER(Job), C.c_int, C.POINTER(C.c_void_p)]
    L.mcfl_batch_run.argtypes = [C.c_void_p, C.POINTER(Job), C.POINTER(C.c_double)]
    L.mcfl_batch_destroy.argtypes = [C.c_void_p]
    L.mcfl_batch_autotune.argtypes = [C.c_void_p, C.c_double, C.POINTER(Tune)]
    L.mcfl_batch_configure.argtypes = [C.c_void_p, C.POINTER(Tune)]
    L.mcfl_batch_destroy.restype = None
    L.mcfl_run_batch.argtypes = [C.c_void_p, SP, C.c_int, C.POINTER(Job), C.c_int, C.POINTER(C.c_double)]
    L.mcfl_server_create.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(C.c_void_p)]
    L.mcfl_server_add_medium.argtypes = [C.c_void_p, C.POINTER(Sim), C.POINTER(C.c_int32)]
    L.mcfl_server_submit.argtypes = [C.c_void_p, C.POINTER(Request), C.POINTER(C.c_int64)]
    L.mcfl_server_step.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.c_int32, C.POINTER(C.c_int32),
                                   C.POINTER(C.c_double)]
    L.mcfl_server_collect.argtypes = [C.c_void_p, C.c_int64, C.POINTER(RequestResult)]
    L.mcfl_server_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_uint64)]
    L.mcfl_server_counters.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.mcfl_server_destroy.argtypes = [C.c_void_p]
    L.mcfl_server_destroy.restype = None
    L._engine = True


EXPORTS += ["mcfl_batch_create", "mcfl_batch_run", "mcfl_batch_destroy", "mcfl_batch_autotune", "mcfl_batch_configure",
            "mcfl_run_batch",
            "mcfl_server_create",
            "mcfl_server_add_medium", "mcfl_server_submit", "mcfl_server_step", "mcfl_server_collect",
            "mcfl_server_stats", "mcfl_server_counters", "mcfl_server_destroy"]


class BatchJob:
    """One simulation of a batch: source, budget, seed, tallies; outputs as numpy arrays."""

    def __init__(self, medium, photons, seed, source=(0.0, 0.0, 0.0), src_kind=SOURCE_PENCIL, src_r=0.0,
                 tallies=0, Fx=None):
        self.medium, self.photons, self.seed = medium, int(photons), int(seed)
        self.source, self.src_kind, self.src_r, self.tallies = source, src_kind, src_r, tallies
        self.Fx = None if Fx is None else np.ascontiguousarray(Fx, np.float64)

    def make(self, sims):
        s = sims[self.medium]
        j = Job()
        j.medium, j.src_kind = self.medium, self.src_kind
        j.xi, j.yi, j.zi = self.source
        j.src_r, j.photons, j.seed, j.tallies = self.src_r, self.photons, self.seed, self.tallies
        j.Fx = _p(self.Fx, F64P)
        self.Rd = np.zeros(s.npix, np.uint64)
        self.Tt = np.zeros(s.npix, np.uint64)
        self.fhd = np.zeros(s.nvox, np.uint64) if self.tallies & TALLY_FHD else None
        self.A_xyz = np.zeros(s.nvox, np.uint64) if self.tallies & TALLY_AXYZ else None
        self.Rd2 = np.zeros(s.npix, np.uint64) if self.tallies & TALLY_SQ else None
        self.Tt2 = np.zeros(s.npix, np.uint64) if self.tallies & TALLY_SQ else None
        nt = s.s.det.x_temp_numdets * s.s.det.y_temp_numdets * s.s.det.temp_bins
        self.time = np.zeros(max(nt, 1), np.uint64) if self.tallies & TALLY_TEMP else None
        j.Rd_xy, j.Tt_xy, j.fhd, j.A_xyz = _p(self.Rd, U64P), _p(self.Tt, U64P), _p(self.fhd, U64P), _p(self.A_xyz, U64P)
        j.Rd2_xy, j.Tt2_xy, j.time_xyt = _p(self.Rd2, U64P), _p(self.Tt2, U64P), _p(self.time, U64P)
        return j


class Tune(C.Structure):
    _fields_ = [("tile_side", C.c_int32), ("tile_voxels", C.c_uint32), ("refill_min", C.c_uint32),
                ("chunk_max", C.c_uint32), ("trial_ms", C.c_double), ("default_ms", C.c_double),
                ("trials", C.c_int32), ("wg", C.c_uint32)]


class Batch:
    """mcfl_batch: media + jobs captured once in a HIP graph, replayed by run()."""

    def __init__(self, ctx, sims, jobs):
        L = lib()
        _bind_engine(L)
        self.sims, self.jobs = list(sims), list(jobs)
        self._sarr = (C.POINTER(Sim) * len(self.sims))(*[C.pointer(s.s) for s in self.sims])
        self._jarr = (Job * len(self.jobs))(*[j.make(self.sims) for j in self.jobs])
        self.h = C.c_void_p()
        check(L.mcfl_batch_create(ctx.h, self._sarr, len(self.sims), self._jarr, len(self.jobs), C.byref(self.h)),
              "mcfl_batch_create")
        ctx._children.add(self)

    def run(self, seeds=None):
        if seeds is not None:
            for j, s in zip(self._jarr, seeds):
                j.seed = int(s)
        ms = C.c_double()
        check(lib().mcfl_batch_run(self.h, self._jarr, C.byref(ms)), "mcfl_batch_run")
        for j, jj in zip(self.jobs, self._jarr):
            j.totals = {n: int(jj.totals[i]) for i, n in enumerate(TOT_NAMES)}
            j.fl_unit = jj.fl_unit
        return ms.value

    def autotune(self, fraction=1.0 / 16):
        """mcfl_batch_autotune: pick the launch settings on a scaled-down copy of the batch."""
        t = Tune()
        check(lib().mcfl_batch_autotune(self.h, float(fraction), C.byref(t)), "mcfl_batch_autotune")
        return {n: getattr(t, n) for n, _ in Tune._fields_}

    def configure(self, wg=0, tile_side=0, refill_min=4, chunk_max=1024):
        """mcfl_batch_configure: set the launch settings (workgroup size, fhd tile side, refill
        threshold, chunk cap) and re-capture the graph; results do not depend on them."""
        t = Tune(tile_side=int(tile_side), refill_min=int(refill_min), chunk_max=int(chunk_max), wg=int(wg))
        check(lib().mcfl_batch_configure(self.h, C.byref(t)), "mcfl_batch_configure")

    def close(self):
        if self.h:
            lib().mcfl_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """mcfl_server: continuous batching of forward-simulation requests."""

    def __init__(self, ctx, mem_budget=8 << 30, epoch_photons=2_000_000):
        L = lib()
        _bind_engine(L)
        self.h = C.c_void_p()
        check(L.mcfl_server_create(ctx.h, int(mem_budget), int(epoch_photons), C.byref(self.h)), "mcfl_server_create")
        ctx._children.add(self)
        self.media = []

    def add_medium(self, sim):
        mid = C.c_int32()
        check(lib().mcfl_server_add_medium(self.h, C.byref(sim.s), C.byref(mid)), "add_medium")
        self.media.append(sim)
        return mid.value

    def submit(self, medium, photons, seed, source=(0.0, 0.0, 0.0), src_kind=SOURCE_PENCIL, src_r=0.0,
               fluorescence=False):
        r = Request()
        r.medium, r.src_kind = medium, src_kind
        r.xi, r.yi, r.zi = source
        r.src_r, r.photons, r.seed, r.fluorescence = src_r, int(photons), int(seed), int(bool(fluorescence))
        rid = C.c_int64()
        check(lib().mcfl_server_submit(self.h, C.byref(r), C.byref(rid)), "submit")
        return rid.value

    def step(self, max_done=256):
        ids = (C.c_int64 * max_done)()
        nd, ms = C.c_int32(), C.c_double()
        check(lib().mcfl_server_step(self.h, ids, max_done, C.byref(nd), C.byref(ms)), "step")
        return [ids[i] for i in range(min(nd.value, max_done))], ms.value

    def collect(self, rid, medium):
        s = self.media[medium]
        out = {"Rd": np.zeros(s.npix, np.uint64), "Tt": np.zeros(s.npix, np.uint64),
               "FlR": np.zeros(s.npix, np.uint64), "FlT": np.zeros(s.npix, np.uint64)}
        r = RequestResult()
        r.Rd_xy, r.Tt_xy = _p(out["Rd"], U64P), _p(out["Tt"], U64P)
        r.FlR_raw, r.FlT_raw = _p(out["FlR"], U64P), _p(out["FlT"], U64P)
        check(lib().mcfl_server_collect(self.h, rid, C.byref(r)), "collect")
        out["fl_unit"] = r.fl_unit
        out["totals"] = {n: int(r.totals[i]) for i, n in enumerate(TOT_NAMES)}
        out["fl_totals"] = {n: int(r.fl_totals[i]) for i, n in enumerate(TOT_NAMES)}
        out["epochs"] = int(r.epochs)
        return out

    def stats(self):
        q, a, b = C.c_int32(), C.c_int32(), C.c_uint64()
        check(lib().mcfl_server_stats(self.h, C.byref(q), C.byref(a), C.byref(b)), "stats")
        return {"queued": q.value, "active": a.value, "bytes": b.value}

    def counters(self):
        """Cumulative photons launched, epochs, kernel ms and host ms spent in steps."""
        p, e, k, h = C.c_uint64(), C.c_uint64(), C.c_double(), C.c_double()
        check(lib().mcfl_server_counters(self.h, C.byref(p), C.byref(e), C.byref(k), C.byref(h)), "counters")
        return {"photons": p.value, "epochs": e.value, "kernel_ms": k.value, "host_ms": h.value}

    def close(self):
        if self.h:
            lib().mcfl_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
