"""Pins of the CPU oracle and of the shared FP32 math / MWC headers.

Each check comes from outside the code it tests: libm in double precision,
the MWC closed form (n-th state = a^n x0 mod a*2^32-1, Python big integers),
textbook Fresnel, Henyey-Greenstein moments, the reference notebook's
back-step known answers (FoundBoundaryPython.ipynb), van de Hulst's published
slab values, Chandrasekhar's H-function for a semi-infinite isotropic medium,
the sources' closed-form distributions (Kolmogorov-Smirnov), the roulette's
exact survival boundary and expectation, the glass march's geometry and an
emission loop whose images are computed by hand.
"""
import math
import pathlib

import numpy as np
import pytest

import _oracle as O
import cudamcfl_amd as M

FLT_MAX = float(np.finfo(np.float32).max)
GOLDEN = pathlib.Path(__file__).parent / "golden" / "ref_inputs"


@pytest.fixture(scope="module")
def a_tab():
    return M.multipliers(32768)


# ------------------------------------------------------------ math header

def test_log_error_bound():
    rng = np.random.default_rng(0)
    u = np.concatenate([rng.uniform(0, 1, 200000), np.geomspace(2.0 ** -24, 1.0, 50000),
                        1.0 - np.arange(1, 20000) * 2.0 ** -24]).astype(np.float32)
    u = u[u > 0]
    got = O.math_eval(0, u).astype(np.float64)
    ref = np.log(u.astype(np.float64))
    ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    err = np.abs(got - ref)
    nz = ref != 0
    assert np.all(err[nz] <= 2.0 * ulp[nz]), f"max ulp error {np.max(err[nz] / ulp[nz])}"
    assert O.math_eval(0, np.array([1.0], np.float32))[0] == 0.0


def test_sincos_turn_error_bound():
    rng = np.random.default_rng(1)
    t = np.concatenate([rng.uniform(0, 1, 200000), np.linspace(0, 1, 4097)]).astype(np.float32)
    s = O.math_eval(1, t).astype(np.float64)
    c = O.math_eval(2, t).astype(np.float64)
    ang = 2 * np.pi * t.astype(np.float64)
    assert np.max(np.abs(s - np.sin(ang))) < 1.5e-7
    assert np.max(np.abs(c - np.cos(ang))) < 1.5e-7
    # quadrant points are exact
    q = np.array([0, 0.25, 0.5, 0.75], np.float32)
    assert np.array_equal(np.abs(O.math_eval(1, q)), [0, 1, 0, 1])
    assert np.array_equal(np.abs(O.math_eval(2, q)), [1, 0, 1, 0])


def test_rsqrt_error_bound():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(6e-8, 1.0, 300000), rng.uniform(0.9, 1.1, 100000),
                        np.geomspace(1e-30, 1e30, 20000)]).astype(np.float32)
    got = O.math_eval(4, x).astype(np.float64)
    ref = 1.0 / np.sqrt(x.astype(np.float64))
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(got - ref) / ulp) <= 2.0


def test_sqrt_and_renormalisation_error_bounds():
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.uniform(1.2e-7, 1.0, 300000), np.geomspace(1.2e-38, 1e30, 20000)]).astype(np.float32)
    got = O.math_eval(5, x).astype(np.float64)
    ref = np.sqrt(x.astype(np.float64))
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(got - ref) / ulp) <= 2.0
    assert O.math_eval(5, np.zeros(1, np.float32))[0] == 0.0
    # near 1 the one-FMA expansion, elsewhere the full root: both within 1 ulp
    y = np.concatenate([1 + rng.uniform(-2 ** -12, 2 ** -12, 200000), rng.uniform(0.5, 2.0, 50000)]).astype(np.float32)
    got = O.math_eval(6, y).astype(np.float64)
    ref = 1.0 / np.sqrt(y.astype(np.float64))
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(got - ref) / ulp) <= 2.0
    near = np.abs(y.astype(np.float64) - 1) < 2 ** -12
    assert np.max(np.abs(got - ref)[near] / ulp[near]) <= 1.0


def test_tof_quotient_is_the_ieee_division():
    """The kernel's time-of-flight increment q / 2.99792e-5 (a product and one
    FMA correction) equals the IEEE double quotient of the reference's
    expression (CUDAMCFLtransport.cu:316) for every non-negative finite float q,
    and for a stretch of negative ones (the formula is odd); the kernel's
    increment (a 32-bit conversion below 2^32) truncates it like the general
    conversion for all of them."""
    assert O.tof_div_mismatches(0, 0x7F800000) == 0
    assert O.tof_div_mismatches(0x80000000, 0x80000000 + (1 << 26)) == 0
    assert O.tof_div_mismatches(0xC0000000, 0xC0000000 + (1 << 26)) == 0


def test_pixel_quotient_as_double_product():
    """The kernel computes the exit tally's pixel quotient q / d
    (CUDAMCFLtransport.cu:374-375,401-402) as the double product with the double
    reciprocal of d, rounded to float.  Its error (2^-52 relative) is below the
    2^-49 relative distance a float quotient keeps from every float midpoint
    it does not equal, so the float is the IEEE quotient: checked on 4M random
    pairs and on 1M quotients built to fall within an ulp of a midpoint."""
    rng = np.random.default_rng(3)
    n = 1 << 22
    d = (10.0 ** rng.uniform(-4, 1, n)).astype(np.float32)
    q = (d * rng.uniform(0, 2000, n)).astype(np.float32)
    got = (q.astype(np.float64) * (1.0 / d.astype(np.float64))).astype(np.float32)
    assert np.array_equal(got, q / d)
    # near midpoints: m = (2k+1) 2^(e-24), q = float(m d) puts q / d next to m
    m = n // 4
    dd = (10.0 ** rng.uniform(-3, 0, m)).astype(np.float32)
    k = rng.integers(1 << 23, 1 << 24, m).astype(np.float64)
    e = rng.integers(-2, 12, m).astype(np.float64)
    mid = (2 * k + 1) * 2.0 ** (e - 24)
    qq = (mid * dd.astype(np.float64)).astype(np.float32)
    got = (qq.astype(np.float64) * (1.0 / dd.astype(np.float64))).astype(np.float32)
    assert np.array_equal(got, qq / dd)
    # the planted mistake: the float reciprocal instead of the double one differs
    bad = (qq * (np.float32(1) / dd)).astype(np.float32)
    assert not np.array_equal(bad, qq / dd)


def test_u32_to_float_toward_zero():
    rng = np.random.default_rng(2)
    u = np.concatenate([rng.integers(0, 2 ** 32, 100000, dtype=np.uint64),
                        [0, 1, 2 ** 24 - 1, 2 ** 24, 2 ** 24 + 1, 2 ** 32 - 1, 2 ** 31 + 255]]).astype(np.uint32)
    got = O.math_eval(3, u.view(np.float32))
    f = u.astype(np.float32)  # nearest
    over = f.astype(np.float64) > u.astype(np.float64)
    f[over] = np.nextafter(f[over], np.float32(0))
    assert np.array_equal(got, f)


# ------------------------------------------------------------- generator

def test_mwc_closed_form():
    a = 4294966893
    p = a * 2 ** 32 - 1
    x0 = (123456 << 32) | 987654321
    states, co = O.mwc_draws(x0, a, 5000)
    for n in (1, 2, 3, 17, 1000, 4999):
        assert int(states[n - 1]) == pow(a, n, p) * x0 % p
    lo = (states & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    f = lo.astype(np.float32)
    over = f.astype(np.float64) > lo.astype(np.float64)
    f[over] = np.nextafter(f[over], np.float32(0))
    assert np.array_equal(co, f * np.float32(2.0 ** -32))
    assert co.min() >= 0 and co.max() < 1


def test_mwc_long_jump():
    a = 4294967118
    p = a * 2 ** 32 - 1
    x0 = 77 << 32 | 5
    states, _ = O.mwc_draws(x0, a, 1_000_000)
    assert int(states[-1]) == pow(a, 1_000_000, p) * x0 % p


def test_primes_table_rule():
    path = M.PRIMES_PATH
    col = np.loadtxt(path, dtype=np.uint64, usecols=0)
    assert len(col) == 32769
    assert np.all(np.diff(col.astype(np.int64)) < 0), "descending order"

    def is_prime(n):
        if n < 2:
            return False
        for sp in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
            if n % sp == 0:
                return n == sp
        d, s = n - 1, 0
        while d % 2 == 0:
            d //= 2
            s += 1
        for b in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
            x = pow(b, d, n)
            if x in (1, n - 1):
                continue
            for _ in range(s - 1):
                x = x * x % n
                if x == n - 1:
                    break
            else:
                return False
        return True

    idx = list(range(0, 20)) + list(range(len(col) - 20, len(col))) + list(range(100, len(col), 997))
    for i in idx:
        a = int(col[i])
        assert is_prime(a * 2 ** 32 - 1) and is_prime(a * 2 ** 31 - 1), a
    # no qualifying multiplier was skipped between the first two lines
    a0, a1 = int(col[0]), int(col[1])
    for a in range(a1 + 1, a0):
        assert not (is_prime(a * 2 ** 31 - 1) and is_prime(a * 2 ** 32 - 1))
    for a in range(a0 + 1, 2 ** 32):
        assert not (is_prime(a * 2 ** 31 - 1) and is_prime(a * 2 ** 32 - 1))


def test_init_rng_matches_seeding_closed_form():
    n = 300
    seed = (4567 << 32) | 891011
    rc, x, a = O.init_rng(n, M.PRIMES_PATH, seed)
    assert rc == 0
    col = np.loadtxt(M.PRIMES_PATH, dtype=np.uint64, usecols=0, max_rows=n + 1)
    begin = int(col[0])
    pb = begin * 2 ** 32 - 1
    assert np.array_equal(a, col[1:].astype(np.uint32))
    k = 0
    for i in range(n):
        # two seeding draws per lane unless a state was rejected (then more)
        while True:
            u1 = (pow(begin, k + 1, pb) * seed % pb) & 0xFFFFFFFF
            u2 = (pow(begin, k + 2, pb) * seed % pb) & 0xFFFFFFFF
            k += 2
            c = math.floor((float(u1) / 4294967296.0) * int(a[i]))
            xi = (c << 32) + u2
            if not (xi == 0 or c >= int(a[i]) - 1 or u2 >= 0xFFFFFFFF):
                break
        assert int(x[i]) == xi, i


def test_init_rng_product_equals_oracle():
    for seed in (1, 2 ** 40 + 12345, 0xFFFFFFF0):
        x, a = M.init_rng(32768, seed=seed)
        rc, xo, ao = O.init_rng(32768, M.PRIMES_PATH, seed)
        assert rc == 0 and np.array_equal(x, xo) and np.array_equal(a, ao)


@pytest.mark.parametrize("seed", [0, (2 ** 32 - 1), (4294967117 << 32) | 5])
def test_init_rng_rejects_invalid_seeds(seed):
    rc, _, _ = O.init_rng(4, M.PRIMES_PATH, seed)
    assert rc == 1
    with pytest.raises(M.McflError):
        M.init_rng(4, seed=seed)


def test_photon_states_valid_and_distinct(a_tab):
    key = M.stream_key(42, 0)
    seen = set()
    for k in range(20000):
        a = int(a_tab[k & 32767])
        s = O.photon_state(key, k, a)
        c, lo = s >> 32, s & 0xFFFFFFFF
        assert c < a - 1 and lo != 0xFFFFFFFF and s != 0
        seen.add(s)
    assert len(seen) == 20000


# ------------------------------------------------------ physics pieces

def fresnel_textbook(n1, n2, ci):
    ti = math.acos(ci)
    st = n1 / n2 * math.sin(ti)
    tt = math.asin(st)
    a, b = ti - tt, ti + tt
    return 0.5 * (math.sin(a) ** 2 / math.sin(b) ** 2 + math.tan(a) ** 2 / math.tan(b) ** 2)


@pytest.mark.parametrize("n1,n2", [(1.4, 1.0), (1.0, 1.4), (1.33, 1.5), (1.5, 1.33)])
def test_fresnel_closed_form(n1, n2):
    lib = O.lib()
    for deg in np.linspace(2, 88, 44):
        ci = math.cos(math.radians(deg))
        if n1 > n2 and math.sin(math.radians(deg)) >= n2 / n1:
            continue
        r = lib.orc_fresnel_test(n1, n2, ci)
        assert abs(r - fresnel_textbook(n1, n2, ci)) < 2e-5 + 1e-4 * fresnel_textbook(n1, n2, ci), deg


def test_reflect_branches():
    import ctypes as C
    lib = O.lib()
    dz, dxs = C.c_float(), C.c_float()
    # total internal reflection: 72.5 deg from the normal, 1.5 -> 1.0
    assert lib.orc_reflect_test(1.5, 1.0, 0.3, 0.99, C.byref(dz), C.byref(dxs)) == 1 and abs(dz.value + 0.3) < 1e-7
    # normal incidence: R = ((1.5-1)/(2.5))^2 = 0.04
    assert lib.orc_reflect_test(1.5, 1.0, 1.0, 0.03, C.byref(dz), C.byref(dxs)) == 1
    assert lib.orc_reflect_test(1.5, 1.0, 1.0, 0.05, C.byref(dz), C.byref(dxs)) == 0 and dz.value == 1.0
    # matched index: always transmitted, direction unchanged
    assert lib.orc_reflect_test(1.4, 1.4, 0.5, 0.0, C.byref(dz), C.byref(dxs)) == 0 and dz.value == 0.5
    # oblique transmission: Snell's law on the tangential component
    r = lib.orc_reflect_test(1.0, 1.5, 0.8, 0.999, C.byref(dz), C.byref(dxs))
    assert r == 0 and abs(dxs.value - 1.0 / 1.5) < 1e-6
    st = math.sqrt(1 - 0.8 ** 2) / 1.5
    assert abs(dz.value - math.sqrt(1 - st * st)) < 1e-6


@pytest.mark.parametrize("g", [0.0, 0.5, 0.9])
def test_spin_henyey_greenstein_moments(g):
    import ctypes as C
    lib = O.lib()
    n = 100000
    x = C.c_uint64((9 << 32) | 1234567)
    a = 4294966893
    din = np.array([0.6, 0.0, 0.8], np.float32)
    dout = np.zeros(3, np.float32)
    mu = np.empty(n)
    norms = np.empty(n)
    for i in range(n):
        lib.orc_spin_test(g, O.ptr(din, O.F32P), C.byref(x), a, O.ptr(dout, O.F32P))
        mu[i] = float(np.dot(din.astype(np.float64), dout.astype(np.float64)))
        norms[i] = float(np.linalg.norm(dout.astype(np.float64)))
    assert np.max(np.abs(norms - 1)) < 1e-6
    assert abs(mu.mean() - g) < 4 * mu.std() / math.sqrt(n) + 1e-4
    m2 = (2 * g * g + 1) / 3
    assert abs((mu ** 2).mean() - m2) < 4 * (mu ** 2).std() / math.sqrt(n) + 1e-4


@pytest.mark.parametrize("g", [0.0, 0.5, 0.9])
def test_spin_equals_the_reference_rotation(g):
    """Spin (CUDAMCFLtransport.cu:585-625) written out here in double
    precision from the reference's own expressions: cos(theta) from the
    Henyey-Greenstein inversion (or 2 xi - 1 from an extra draw when g = 0),
    psi = 2 pi xi with the sine assigned to cosp and the cosine to sinp
    (:603), the normal-incidence branch when 1 - dz^2 rounds to 0 (:607-612)
    and the general rotation otherwise (:613-618), then renormalised.  The
    draws are the generator's (pinned by its closed form).  The oracle's
    FP32 rewrite (one reciprocal root for sint/temp and sint*temp, a
    first-order renormalisation) agrees to 3e-6 + 1e-6/sin(theta) (FP32's
    own conditioning near cos(theta) = +-1) on every component for 4000
    directions, among them dz = +-1 and dz within 2^-12 of +-1.  Swapping
    the sine and cosine back, or the sign of the dz term, fails it."""
    import ctypes as C
    lib = O.lib()
    a = 4294966893
    rng = np.random.default_rng(17)
    dirs = rng.normal(size=(4000, 3))
    dirs[:8] = [[0, 0, 1], [0, 0, -1], [1e-4, 0, 1], [0, -3e-4, -1], [0.6, 0, 0.8], [0, 1, 0], [1, 0, 0],
                [0.3, 0.4, 0]]
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    g32 = float(np.float32(g))
    worst = 0.0
    for i, d in enumerate(dirs.astype(np.float32)):
        x0 = ((i + 3) << 32) | (2654435761 * (i + 1) % 2 ** 32)
        _, draws = O.mwc_draws(x0, a, 3)
        dr = draws.astype(np.float64)
        if g32 != 0.0:
            t = (1 - g32 * g32) / (1 - g32 + 2 * g32 * dr[0])
            cost = (1 + g32 * g32 - t * t) / (2 * g32)
            u = dr[1]
        else:
            cost = 2 * dr[1] - 1
            u = dr[2]
        psi = 2 * math.pi * u
        cosp, sinp = math.sin(psi), math.cos(psi)  # the reference's argument order
        sint = math.sqrt(max(0.0, 1 - cost * cost))
        dx, dy, dz = map(float, d)
        if np.float32(1) - np.float32(dz) * np.float32(dz) <= 0:
            nd = np.array([sint * cosp, sint * sinp, math.copysign(cost, dz * cost)])
        else:
            temp = math.sqrt(1 - dz * dz)
            nd = np.array([sint * (dx * dz * cosp - dy * sinp) / temp + dx * cost,
                           sint * (dy * dz * cosp + dx * sinp) / temp + dy * cost,
                           -sint * cosp * temp + dz * cost])
        nd /= np.linalg.norm(nd)
        xs = C.c_uint64(x0)
        dout = np.zeros(3, np.float32)
        lib.orc_spin_test(g, O.ptr(np.ascontiguousarray(d), O.F32P), C.byref(xs), a, O.ptr(dout, O.F32P))
        # FP32 conditioning: near cos(theta) = +-1 an error e in cos(theta) becomes e / sin(theta) in sin(theta)
        tol = 3e-6 + 1e-6 / max(sint, 1e-6)
        worst = max(worst, float(np.max(np.abs(dout.astype(np.float64) - nd))) / tol)
    assert worst < 1.0, worst


def test_backstep_known_answers():
    import ctypes as C
    lib = O.lib()
    t1, t2 = C.c_float(), C.c_float()
    f = lambda v: np.array(v, np.float32)
    lib.orc_backstep_test(O.ptr(f([0.8, 1.2, 1]), O.F32P), O.ptr(f([0, 0, 1]), O.F32P), O.ptr(f([1, 1, 1]), O.F32P),
                          0.25, C.byref(t1), C.byref(t2))
    assert abs(t1.value + 0.412310562562) < 1e-6 and abs(t2.value - 0.412310562562) < 1e-6
    r = np.float32(0.141421)
    lib.orc_backstep_test(O.ptr(f([1.489575, -0.333571, 1.599638]), O.F32P),
                          O.ptr(f([0.422606, -0.768113, -0.481048]), O.F32P),
                          O.ptr(f([1.4, -0.4, 1.6]), O.F32P), float(r * r), C.byref(t1), C.byref(t2))
    assert abs(t1.value + 0.10092855989933014) < 2e-6 and abs(t2.value - 0.07493677735328674) < 2e-6


# -------------------------------------------------- whole-transport pins

def homogeneous(mua, mus, g, d, gs, n_in=1.0, n_out=1.0):
    nz = int(round(d * gs))
    nx = int(4 * np.float32(d) * gs)
    lab = np.ones((nz, nx, nx), np.uint8)
    f32 = np.float32
    b = [[FLT_MAX, 0, 1, n_out, 0, 0, 0, 0],
         [f32(1) / f32(mua + mus), mua, g, n_in, 0, mua, 0, 1],
         [FLT_MAX, 0, 1, n_out, 0, 0, 0, 0]]
    return O.Medium(lab, b, d, gs)


def test_van_de_hulst_slab(a_tab):
    """mua=10, mus=90 /cm, g=0.75, n=1, d=0.02 cm: Rd = 0.09739, Tt = 0.66096
    (van de Hulst; MCML's validation table)."""
    m = homogeneous(10, 90, 0.75, 0.02, 50)
    N = 300000
    r = O.run_photons(m.sim(src=(1, 0, 0, 0, 0)), a_tab, 2024, 0, N)
    W = N * 0xFFFFFFFF
    t = r.tot
    assert abs(t["R"] / W - 0.09739) < 0.0025
    assert abs(t["T"] / W - 0.66096) < 0.003


def chandrasekhar_rd(albedo, mu0=1.0, n=64):
    x, w = np.polynomial.legendre.leggauss(n)
    mu = 0.5 * (x + 1)
    w = 0.5 * w
    H = np.ones(n)
    for _ in range(500):
        Hn = 1.0 / (math.sqrt(1 - albedo) + 0.5 * albedo * (w * mu * H / (mu[:, None] + mu)).sum(axis=1))
        if np.max(np.abs(Hn - H)) < 1e-14:
            break
        H = Hn
    H0 = 1.0 / (math.sqrt(1 - albedo) + 0.5 * albedo * np.sum(w * mu * H / (mu0 + mu)))
    return 1 - H0 * math.sqrt(1 - albedo)


def test_semi_infinite_isotropic_chandrasekhar(a_tab):
    """Isotropic scattering, albedo 0.9, matched index, normal incidence:
    R = 1 - H(1) sqrt(1 - albedo)."""
    m = homogeneous(10, 90, 0.0, 1.0, 5)
    N = 200000
    r = O.run_photons(m.sim(src=(1, 0, 0, 0, 0)), a_tab, 77, 0, N)
    t = r.tot
    R = t["R"] / (N * 0xFFFFFFFF)
    assert t["T"] == 0
    assert abs(R - chandrasekhar_rd(0.9)) < 0.006, (R, chandrasekhar_rd(0.9))


def test_energy_balance_and_counts(a_tab):
    import media
    sim = media.small_sim(photons=3000)
    o = O.FromProduct(sim, axyz=True)
    r = O.run_photons(o.o, a_tab, 5, 0, 3000)
    t = r.tot
    assert t["photons"] == 3000 and t["aborted"] == 0
    assert t["launched_w"] == 3000 * sim.s.start_weight
    assert t["launched_w"] + t["roul_gain"] - t["roul_loss"] - t["zero_kill"] == t["R"] + t["T"] + t["A"]
    # absorption laterally outside the grid (medium 1, infinite) has no voxel
    assert 0.5 * t["A"] < int(r.axyz.sum()) <= t["A"]
    assert int(r.Rd.sum()) <= t["R"] and int(r.Tt.sum()) <= t["T"]


def test_step_cap_keeps_the_energy_balance(a_tab):
    """With the per-photon step cap at 40 (default 2^28), photons alive after
    their 40th step stop, their weight goes to the aborted total, no photon
    takes more than 40 steps, and the balance still closes; with the cap at 39
    more weight is aborted."""
    import media
    sim = media.small_sim(photons=3000)
    o = O.FromProduct(sim)
    res = {}
    try:
        for cap in (40, 39):
            O.set_max_steps(cap)
            res[cap] = O.run_photons(o.o, a_tab, 5, 0, 3000).tot
    finally:
        O.set_max_steps(0)
    for cap, t in res.items():
        assert t["aborted"] > 0 and t["steps"] <= cap * 3000
        assert t["launched_w"] + t["roul_gain"] - t["roul_loss"] - t["zero_kill"] - t["aborted"] == \
            t["R"] + t["T"] + t["A"]
    assert res[39]["aborted"] > res[40]["aborted"]


def test_thread_count_does_not_change_tallies(a_tab):
    import media
    sim = media.small_sim(photons=2000)
    o = O.FromProduct(sim)
    r1 = O.run_photons(o.o, a_tab, 9, 0, 2000, nthreads=1)
    r8 = O.run_photons(o.o, a_tab, 9, 0, 2000, nthreads=8)
    assert np.array_equal(r1.Rd, r8.Rd) and np.array_equal(r1.fhd, r8.fhd) and r1.tot == r8.tot


def test_lane_and_photon_streams_agree_statistically(a_tab):
    """The reference's per-lane streams and this build's per-photon streams
    sample the same transport: totals agree within Monte Carlo noise."""
    import media
    N = 40000
    sim = media.small_sim(photons=N, temporal=False)
    o = O.FromProduct(sim)
    rp = O.run_photons(o.o, a_tab, 31, 0, N)
    x, a = M.init_rng(1024, seed=99)
    rl = O.run_lanes(o.o, x, a, N)
    for k in ("R", "T", "A"):
        p, l = rp.tot[k] / rp.tot["launched_w"], rl.tot[k] / rl.tot["launched_w"]
        assert abs(p - l) < 5 * math.sqrt(max(p, 1e-3) / N), (k, p, l)


def test_lanes_require_enough_photons(a_tab):
    import media
    sim = media.small_sim(photons=100)
    o = O.FromProduct(sim)
    x, a = M.init_rng(512, seed=3)
    with pytest.raises(ValueError):
        O.run_lanes(o.o, x, a, 100)


# --------------------------------------------- bulk method 1 (layered slab)

FAR = (50.0, 50.0, 50.0, 0.001, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)  # inclusion no photon reaches


def test_layered_forward_beam_exact(a_tab):
    """g = 1 keeps every photon on the axis (cos theta = 1 exactly), and outside
    the inclusion the step is the deterministic mutr (CUDAMCFLtransport.cu:99),
    so every photon follows the same path: steps of mutr clipped at the layer
    bounds, no absorption at the step that changes layer, w -= rn(mua mutr w)
    otherwise.  Walk that path in float32 and compare the totals exactly."""
    f32 = np.float32
    layers = [(1.0, 0.3, 40.0, 0, 0, 0, 0, 1.0, 0.05), (1.0, 0.8, 25.0, 0, 0, 0, 0, 1.0, 0.03)]
    sim = M.Simulation.layered(layers, FAR, grid_size=20, det=(0.05, 0.05, 8, 8, 0.0, 0.0), photons=300)
    o = O.FromProduct(sim)
    r = O.run_photons(o.o, a_tab, M.stream_key(3, 0), 0, 300)
    s = sim.s
    zmin = [f32(s.layers[i].z_min) for i in range(4)]
    zmax = [f32(s.layers[i].z_max) for i in range(4)]
    mutr = [f32(s.layers[i].mutr) for i in range(4)]
    mua_mutr = [f32(f32(s.layers[i].mua) * f32(s.layers[i].mutr)) for i in range(4)]
    z, L, w, steps, absorbed = f32(0), 1, int(s.start_weight), 0, 0
    while True:
        steps += 1
        st = mutr[L]
        nl = L
        if f32(z + st) >= zmax[L]:
            nl += 1
            st = f32(zmax[L] - z)
        z = f32(z + st)
        if nl != L:
            L = nl
            if L > 2:
                break
            continue
        dw = int(np.rint(f32(mua_mutr[L] * f32(w))))
        w -= dw
        absorbed += dw
    assert w > 429497  # the roulette never draws on this path
    t = r.tot
    assert t["T"] == 300 * w and t["R"] == 0
    assert t["A"] == 300 * absorbed and t["steps"] == 300 * steps
    assert t["launched_w"] == 300 * int(s.start_weight)


def test_layered_slab_van_de_hulst(a_tab):
    """An inclusion that contains the whole slab makes method 1 sample its
    steps (:87-91) with the inclusion's optics; matched indices at both
    surfaces: van de Hulst's slab, mua=10, mus=90 /cm, g=0.75, d=0.02 cm,
    Rd = 0.09739, Tt = 0.66096."""
    layers = [(1.0, 1.0, 10.0, 0, 0, 0, 0, 0.5, 0.02)]
    inc = (0.0, 0.0, 0.01, 100.0, 1.0, 10.0, 90.0, 0.0, 0.0, 0.0, 0.0, 0.75)
    sim = M.Simulation.layered(layers, inc, grid_size=50, photons=10)
    o = O.FromProduct(sim)
    N = 300000
    r = O.run_photons(o.o, a_tab, 2025, 0, N)
    W = N * 0xFFFFFFFF
    t = r.tot
    assert abs(t["R"] / W - 0.09739) < 0.0025
    assert abs(t["T"] / W - 0.66096) < 0.003


def test_layered_sample_mci_energy_identity(a_tab):
    """The reference's own method-1 input (index step 1.33/1.0, one inclusion,
    temporal detectors): launched + roulette gain - roulette loss = R + T + A
    exactly in integers, and no weight is dropped at zero absorption."""
    sim = M.Simulation.from_mci(str(GOLDEN / "sample.mci"))
    sim.s.number_of_photons = 3000
    sim.s.fhd_activated = 1
    o = O.FromProduct(sim)
    assert o.o.layered == 1 and o.o.start_layer == 1
    r = O.run_photons(o.o, a_tab, 11, 0, 3000)
    t = r.tot
    assert t["photons"] == 3000 and t["zero_kill"] == 0
    assert t["launched_w"] + t["roul_gain"] - t["roul_loss"] == t["R"] + t["T"] + t["A"]
    assert 0 < t["R"] and 0 < t["A"]
    assert r.fhd.sum() > 0


def test_fluorescence_estimators_agree(a_tab):
    """The reference's emission loop (one simulation per emitting voxel on
    lane streams that continue from voxel to voxel, CUDAMCFLmain.cu:471-594)
    and this build's voxel emission jobs on photon streams estimate the same
    FlR and FlT images: totals agree within Monte Carlo noise over 4 seeds
    each."""
    import media
    lab = media.small_labels(10, 10, 5, 4)
    sim = M.Simulation.voxel(lab, media.MATS3, esp=0.5, grid_size=10, det=(0.1, 0.1, 10, 10, 0.0, 0.0),
                             photons=20000, n_above=1.0, n_below=1.33, fhd=True)
    o = O.FromProduct(sim)
    ex = O.run_photons(o.o, a_tab, M.stream_key(3, 0), 0, 20000)
    Fx = ex.fhd.astype(np.float64) / (4294967295.0 * 20000)
    oe = O.FromProduct(sim, emission=True)
    nv, L = 130, 64
    xy = 100
    dxdy = float(np.float32(0.1)) * float(np.float32(0.1))
    vox, scale, unit = M.fl_plan(sim, Fx, nv)
    ref, ps = [], []
    for seed in range(4):
        x, a = M.init_rng(L, seed=seed + 1)
        FlR, FlT, n = O.fluor_reference(oe.o, Fx, x, a, nv)
        assert n > 0
        ref.append((FlR.sum(), FlT.sum()))
        keys = np.array([M.stream_key(seed + 10, int(v)) for v in vox], np.uint64)
        r = O.run_fluor_jobs(oe.o, a_tab, vox, keys, scale, nv)
        ps.append((r.Rd[:xy].sum() * unit / dxdy, r.Tt[:xy].sum() * unit / dxdy))
    ref, ps = np.array(ref), np.array(ps)
    for k in range(2):
        m1, m2 = ref[:, k].mean(), ps[:, k].mean()
        sig = math.sqrt(ref[:, k].var(ddof=1) / 4 + ps[:, k].var(ddof=1) / 4)
        assert m1 > 0 and abs(m1 - m2) < 5 * sig + 1e-3 * m1, (k, m1, m2, sig)


def test_fixed_point_lane_emission_equals_reference_loop(a_tab):
    """The lane-stream emission loop with fixed-point images (what the GPU runs)
    reproduces the reference's loop with its double accumulation
    (CUDAMCFLmain.cu:552-577): same advanced lane states, images equal to
    1e-9 relative."""
    import media
    lab = media.small_labels(10, 10, 5, 4)
    sim = M.Simulation.voxel(lab, media.MATS3, esp=0.5, grid_size=10, det=(0.1, 0.1, 10, 10, 0.0, 0.0),
                             photons=20000, n_above=1.0, n_below=1.33, fhd=True)
    o = O.FromProduct(sim)
    ex = O.run_photons(o.o, a_tab, M.stream_key(3, 0), 0, 20000)
    Fx = ex.fhd.astype(np.float64) / (4294967295.0 * 20000)
    oe = O.FromProduct(sim, emission=True)
    nv, L = 100, 64
    x0, a = M.init_rng(L, seed=5)
    x1 = x0.copy()
    FlR, FlT, _ = O.fluor_reference(oe.o, Fx, x1, a, nv)
    vox, scale, unit = M.fl_plan(sim, Fx, nv)
    vox, scale = O.reference_order(sim, vox, scale)
    x2 = x0.copy()
    r = O.run_fluor_lanes(oe.o, x2, a, vox, scale, nv)
    assert np.array_equal(x1, x2)
    dxdy = float(np.float32(0.1)) * float(np.float32(0.1))
    for img, raw in ((FlR, r.Rd), (FlT, r.Tt)):
        fx = raw[:100].astype(np.float64) * unit / dxdy
        assert img.sum() > 0
        np.testing.assert_allclose(fx, img, rtol=1e-9, atol=1e-12 * img.max())
    with pytest.raises(ValueError):
        O.run_fluor_lanes(oe.o, x0.copy(), a, vox, scale, L - 1)


# ------------------------------------- sources, roulette, glass march, emission loop

def _ks_uniform(u):
    """Kolmogorov-Smirnov distance of samples u from U[0, 1]."""
    u = np.sort(np.asarray(u, np.float64))
    n = len(u)
    i = np.arange(1, n + 1)
    return max(np.max(i / n - u), np.max(u - (i - 1) / n))


KS_CRIT = 1.95  # D * sqrt(n) above this rejects at the 0.1% level


def _source_medium(gs=10):
    """8x8x6 voxels of 1/gs cm, labels 1 and 2 in a checkerboard of voxels."""
    iz, iy, ix = np.indices((6, 8, 8))
    lab = (1 + (ix + iy + iz) % 2).astype(np.uint8)
    f32 = np.float32
    b = [[FLT_MAX, 0, 1, 1.0, 0, 0, 0, 0], [f32(1) / f32(100.1), 0.1, 0.9, 1.4, 0, 0.1, 0.1, 1],
         [f32(1) / f32(100.5), 0.5, 0.9, 1.4, 0, 0.5, 0.1, 1], [FLT_MAX, 0, 1, 1.0, 0, 0, 0, 0]]
    return O.Medium(lab, b, 6 / gs, gs, halfx=8 / (2 * gs), halfy=8 / (2 * gs))


def _expected_label(m, x, y, z):
    """Label of the voxel holding each point (x fastest, CUDAMCFLtransport.cu:490-503)."""
    ix = np.floor((x + float(m.halfx)) * m.gs).astype(np.int64)
    iy = np.floor((y + float(m.halfy)) * m.gs).astype(np.int64)
    iz = np.floor(z * m.gs).astype(np.int64)
    return m.labels[iz, iy, ix]


def test_pencil_source_radial_law(a_tab):
    """The collimated source of LaunchPhoton (CUDAMCFLtransport.cu:507-543)
    over 10^6 launches: the spot radius follows P(r < R) = 1 - exp(-R^2 /
    0.03^2) and its azimuth is uniform (Kolmogorov-Smirnov against the
    closed forms); every photon starts at z = zi moving along +z with the
    start weight, labelled with the voxel it enters.  A radius
    0.03 * (-ln xi), without the root, fails the radial test (checked by
    planting it in a copy of the oracle)."""
    m = _source_medium()
    xi, yi, zi = np.float32(0.013), np.float32(-0.021), np.float32(0.0)
    sim = m.sim(src=(1, float(xi), float(yi), float(zi), 0.0), start_weight=4123456789)
    n = 1_000_000
    pd, w, lab = O.launch_sample(sim, 1, a_tab, M.stream_key(41, 0), n)
    x, y = pd[:, 0].astype(np.float64) - float(xi), pd[:, 1].astype(np.float64) - float(yi)
    r2 = x * x + y * y
    assert _ks_uniform(1.0 - np.exp(-r2 / 0.03 ** 2)) * math.sqrt(n) < KS_CRIT
    phi = np.mod(np.arctan2(y, x), 2 * np.pi) / (2 * np.pi)
    assert _ks_uniform(phi) * math.sqrt(n) < KS_CRIT
    assert np.all(pd[:, 2] == zi) and np.all(pd[:, 3] == 0) and np.all(pd[:, 4] == 0) and np.all(pd[:, 5] == 1)
    assert np.all(w == 4123456789)
    assert np.array_equal(lab, _expected_label(m, pd[:, 0].astype(np.float64), pd[:, 1].astype(np.float64),
                                               pd[:, 2].astype(np.float64)))
    assert len(np.unique(lab)) == 2  # the spot covers voxels of both labels


def test_isotropic_source_distributions(a_tab):
    """The isotropic source of LaunchPhoton (CUDAMCFLtransport.cu:466-503) at
    a voxel corner, over 10^6 launches: the position is uniform in
    (xi, xi + 0.999999/gs] on each axis and never leaves the voxel, cos(theta)
    is uniform on [-1, 1] and the azimuth on [0, 2 pi) (Kolmogorov-Smirnov),
    the direction is a unit vector, the weight 0xFFFFFFFF and the label that
    of the corner's voxel.  cos(theta) = 1 - xi instead of 1 - 2 xi fails the
    cos(theta) test, and a position step 1/gs instead of 0.999999/gs fails
    the bounds (both checked by planting them in a copy of the oracle)."""
    m = _source_medium()
    gs = m.gs
    corner = np.array([-0.3, 0.2, 0.1], np.float32)  # voxel (1, 6, 1)
    sim = m.sim(src=(0, float(corner[0]), float(corner[1]), float(corner[2]), 0.0))
    n = 1_000_000
    pd, w, lab = O.launch_sample(sim, 0, a_tab, M.stream_key(42, 0), n)
    step = float(np.float32(np.float32(0.999999) / np.float32(gs)))
    for k in range(3):
        c = float(corner[k])
        p = pd[:, k].astype(np.float64)
        assert p.min() >= c and p.max() <= float(np.float32(c + step)) and p.max() < c + 1.0 / gs
        assert _ks_uniform((p - c) / step) * math.sqrt(n) < KS_CRIT
    d = pd[:, 3:6].astype(np.float64)
    assert np.max(np.abs(np.sum(d * d, axis=1) - 1.0)) < 1e-6
    assert _ks_uniform((d[:, 2] + 1.0) / 2.0) * math.sqrt(n) < KS_CRIT
    psi = np.mod(np.arctan2(d[:, 1], d[:, 0]), 2 * np.pi) / (2 * np.pi)
    assert _ks_uniform(psi) * math.sqrt(n) < KS_CRIT
    assert np.all(w == 0xFFFFFFFF)
    assert np.all(lab == m.labels[1, 6, 1])


def _state_drawing(u, a):
    """A generator state whose next [0,1) draw comes from the 32-bit value u:
    carry 0 and low word u * a^-1 mod 2^32."""
    return (u * pow(a, -1, 2 ** 32)) % 2 ** 32


def test_roulette_rule_and_expectation(a_tab):
    """PhotonSurvive (CUDAMCFLtransport.cu:850-864): weights above 429497
    live without a draw, weight 0 dies without one; weights 1..429497 draw
    once and live with probability 0.1f, their weight becoming w / 0.1
    rounded, else die.  The survival boundary in the draw's 32-bit value is
    found by bisection and equals 0.1f * 2^32 = 429496736 (0.1f is
    13421773 * 2^-27), so E[w'] = w to within the rounding of w / 0.1 for
    every weight tried; over 10^6 photon streams the mean surviving weight is
    w within 4 sigma.  A rule `draw <= 0.1` moves the boundary by one float
    step and fails; a weight w / 0.2 fails the expectation (both planted in
    a copy of the oracle)."""
    a = int(a_tab[0])
    # no draw: above the threshold, or weight 0
    for w, lives in ((429498, 1), (0xFFFFFFFF, 1), (0, 0)):
        x0 = (123 << 32) | 456
        r, wo, x1, g, lo = O.survive(w, x0, a)
        assert (r, x1, g, lo) == (lives, x0, 0, 0) and wo == w
    # the boundary of the draw (a u32 whose conversion toward zero times 2^-32 is the draw)
    lo_u, hi_u = 0, 2 ** 32 - 1
    assert O.survive(1000, _state_drawing(lo_u, a), a)[0] == 1 and O.survive(1000, _state_drawing(hi_u, a), a)[0] == 0
    while hi_u - lo_u > 1:
        mid = (lo_u + hi_u) // 2
        if O.survive(1000, _state_drawing(mid, a), a)[0]:
            lo_u = mid
        else:
            hi_u = mid
    assert hi_u == 13421773 * 32 == 429496736
    p_live = hi_u / 2 ** 32
    for w in [1, 2, 3, 7, 1000, 12345, 65535, 99999, 300000, 429496, 429497]:
        x0 = _state_drawing(hi_u - 1, a)
        r, wo, x1, g, lo = O.survive(w, x0, a)
        assert r == 1 and abs(wo - 10 * w) <= 1 and g == wo - w and lo == 0 and x1 != x0
        assert abs(p_live * wo - w) <= 0.1 + 2e-8 * w
        r, wo, x1, g, lo = O.survive(w, _state_drawing(hi_u, a), a)
        assert r == 0 and wo == 0 and lo == w and g == 0
    n, w = 1_000_000, 300000
    lived, total = O.survive_batch(w, a_tab, M.stream_key(43, 0), n)
    sigma = 10 * w * math.sqrt(0.1 * 0.9 / n)
    assert abs(total / n - w) < 4 * sigma
    assert abs(lived / n - 0.1) < 4 * math.sqrt(0.09 / n)


def _glass_medium():
    """8x8x10 voxels of 1 mm: scattering label 1 with a glass layer (label 2,
    mus = 0) at voxel depths 2..6."""
    lab = np.ones((10, 8, 8), np.uint8)
    lab[2:7] = 2
    f32 = np.float32
    b = [[FLT_MAX, 0, 1, 1.0, 0, 0, 0, 0], [f32(1) / f32(100.1), 0.1, 0.9, 1.4, 0, 0.1, 0.1, 1],
         [FLT_MAX, 0.0, 1.0, 1.5, 0, 0, 0, 1], [FLT_MAX, 0, 1, 1.0, 0, 0, 0, 0]]
    return O.Medium(lab, b, 1.0, 10, halfx=0.4, halfy=0.4)


def _march_geometry(m, p0, d, h):
    """The half-voxel march of MoveToFirstBoundary in double precision
    (CUDAMCFLtransport.cu:717-756): position and voxel (ix, iy, iz) of the
    first step that lands in a voxel of another label, and the step count."""
    p = np.array(p0, np.float64)
    for k in range(1, 1000):
        p = p + h * d
        nxt = p + h * d
        assert 0.0 <= nxt[2] <= float(m.esp) and abs(nxt[0]) < float(m.halfx) and abs(nxt[1]) < float(m.halfy)
        v = np.floor((p + [float(m.halfx), float(m.halfy), 0.0]) * m.gs).astype(int)
        frac = (p + [float(m.halfx), float(m.halfy), 0.0]) * m.gs - v
        assert np.all(np.minimum(frac, 1 - frac) > 1e-4)  # no march point near a voxel face
        if m.labels[v[2], v[1], v[0]] != 2:
            return p, v, k
    raise AssertionError("no boundary")


def test_glass_march_lands_on_the_backoff_sphere():
    """MoveToFirstBoundary (CUDAMCFLtransport.cu:702-847) through a glass
    layer: the photon marches in half-voxel steps (the time of flight grows by
    trunc(0.05 cm / c) per step) to the first voxel of another label, whose
    label it returns, then backs off along its ray onto the sphere of radius
    0.707107/gs around that voxel's centre.  The expected voxel, step count
    and the ray's two crossings of that sphere are computed here from the
    geometry in double precision; the landing point is one of the crossings.
    A march of whole voxels fails the time of flight; a sphere of radius
    0.5/gs (the reference comment's value) misses both crossings (both
    planted in a copy of the oracle)."""
    m = _glass_medium()
    sim = m.sim()
    c_cmfs = 2.99792e-5
    h = float(np.float32(1.0) / (np.float32(2.0) * np.float32(10.0)))
    cases = [((0.013, -0.027, 0.234), (0.3, -0.2, 0.93), 1),   # down into the scattering layer below
             ((0.052, 0.031, 0.612), (-0.25, 0.4, -0.88), 1),  # up into the layer above
             ((-0.151, 0.143, 0.466), (0.55, 0.45, 0.704), 1)]  # steeply, through several columns
    for p0, d0, want in cases:
        d = np.array(d0, np.float32)
        d = (d / np.float32(np.sqrt(np.sum(d.astype(np.float64) ** 2)))).astype(np.float32)
        p0 = np.array(p0, np.float32)
        lab, out, tof = O.march(sim, p0, d, 2)
        pk, v, k = _march_geometry(m, p0.astype(np.float64), d.astype(np.float64), h)
        assert lab == want == m.labels[v[2], v[1], v[0]]
        c = (v + 0.5) / m.gs - [float(m.halfx), float(m.halfy), 0.0]
        r = 0.707107 / m.gs
        dd = d.astype(np.float64)
        q = pk - c
        b = dd @ q
        disc = b * b - (dd @ dd) * (q @ q - r * r)
        assert disc > 0
        s = [(-b - math.sqrt(disc)) / (dd @ dd), (-b + math.sqrt(disc)) / (dd @ dd)]
        cands = [pk + si * dd for si in s]
        dist = [np.max(np.abs(out - cc)) for cc in cands]
        j = int(np.argmin(dist))
        assert dist[j] < 2e-6, (dist, out, cands)
        back = -s[j]  # the reference moves by p - t d
        step_t = int(h / c_cmfs)
        assert abs(tof - (k * step_t - max(0, int(back / c_cmfs)))) <= 1, (tof, k)


def test_glass_march_bound_ends_only_endless_marches():
    """The march's pass bound (4 max_s gs + 64 passes).  A photon inside an
    exit-labelled region (label 0 above, n_bulks+1 below) that marches to the
    slab's surface never sees its label change, and the reference's loop
    (CUDAMCFLtransport.cu:717-756) steps by -z/dz = 0 at the surface for
    ever: the bounded march returns the exit label with the photon on the
    surface.  A glass march that leaves the grid sideways in a glass label 1
    (the label of the space beside the grid, :741-744) runs until the float
    sum of its 0.05 cm steps reaches max_s = 100 cm, 2000 passes, as in the
    reference: the bound must not cut it short (a bound of max_s gs passes
    fails the time of flight)."""
    f32 = np.float32
    glass = [FLT_MAX, 0.0, 1.0, 1.5, 0, 0, 0, 1]
    lab = np.full((10, 8, 8), 2, np.uint8)
    lab[:4] = 0  # exit label 0 in the top 4 voxel layers, 3 (= n_bulks+1) in the bottom 3
    lab[7:] = 3
    b = [[FLT_MAX, 0, 1, 1.0, 0, 0, 0, 0], glass, [f32(1) / f32(100.1), 0.1, 0.9, 1.4, 0, 0.1, 0.1, 1],
         [FLT_MAX, 0, 1, 1.0, 0, 0, 0, 0]]
    m = O.Medium(lab, b, 1.0, 10, halfx=0.4, halfy=0.4)
    sim = m.sim()
    for p0, d0, old, z_end in (((0.01, 0.02, 0.25), (0.1, -0.2, -0.97), 0, 0.0),
                               ((-0.03, 0.01, 0.85), (0.2, 0.1, 0.97), 3, 1.0)):
        d = np.array(d0, np.float32)
        d = (d / np.float32(np.sqrt(np.sum(d.astype(np.float64) ** 2)))).astype(np.float32)
        lab_out, out, _ = O.march(sim, np.array(p0, np.float32), d, old)
        assert lab_out == old
        assert abs(float(out[2]) - z_end) < 1e-6
        t = (z_end - p0[2]) / float(d[2])  # the ray's crossing of the surface
        assert np.max(np.abs(out[:2] - (np.array(p0[:2]) + t * d[:2].astype(np.float64)))) < 1e-5
    # the long sideways march in glass label 1
    lab1 = np.ones((10, 8, 8), np.uint8)
    m1 = O.Medium(lab1, [b[0], glass, b[3]], 1.0, 10, halfx=0.4, halfy=0.4)
    sim1 = m1.sim()
    d = np.array([0.6, 0.8, 1e-4], np.float32)
    lab_out, out, tof = O.march(sim1, np.array([0.0, 0.0, 0.5], np.float32), d, 1)
    h = f32(1) / (f32(2) * f32(10))
    total, k = f32(0), 0
    while total < f32(100):
        total = f32(total + h)
        k += 1
    assert lab_out == 1 and 1990 <= k <= 2010
    assert tof == k * int(float(h) / 2.99792e-5)
    assert abs(float(out[0]) - k * float(h) * 0.6) < 1e-3 * k * float(h)


def test_emission_loop_images_by_hand(a_tab):
    """The reference's emission loop (CUDAMCFLmain.cu:471-594) on a 3x1x1
    medium whose photons all leave on their first step (mean free path
    10^7 cm in a 1 mm slab, matched indices, one 10^6 cm detector pixel):
    each emitting voxel adds eY (1 - albedof) Fx R / (N (2^32 - 1) dx dy) to
    FlR (likewise T to FlT, :552-577), and with every photon leaving through
    the top or the bottom at full weight the two images sum to
    sum_v eY_v (1 - albedof_v) Fx_v / (dx dy), computed here by hand; the
    voxel of a material with albedof = 1 emits nothing but counts one
    photon (:558-561); each emitting voxel's photons split evenly between
    top and bottom (isotropic source) within 4 sigma.  Dropping (1 - albedof)
    or dividing by N twice fails the sum (both planted in a copy of the
    oracle)."""
    lab = np.array([[[1, 2, 3]]], np.uint8)
    f32 = np.float32
    mutr = float(f32(1) / f32(1e-7))
    b = [[FLT_MAX, 0, 1, 1.0, 0, 0, 0, 0],
         [mutr, 0.0, 0.9, 1.0, 1.0, 0.0, 0.1, 0.5],
         [mutr, 0.0, 0.9, 1.0, 1.0, 0.0, 0.3, 0.25],
         [mutr, 0.0, 0.9, 1.0, 1.0, 0.0, 0.7, 1.0],
         [FLT_MAX, 0, 1, 1.0, 0, 0, 0, 0]]
    m = O.Medium(lab, b, 0.1, 10, halfx=0.15, halfy=0.05)
    sim = m.sim(det=(1e6, 1e6, 0.0, 0.0, 1, 1), src=(0, 0.0, 0.0, 0.0, 0.0))
    Fx = np.array([0.37, 0.81, 0.55])
    nv, L = 4096, 64
    x, a = M.init_rng(L, seed=8)
    FlR, FlT, n = O.fluor_reference(sim, Fx, x, a, nv)
    assert n == 2 * nv + 1
    dxdy = float(f32(1e6)) ** 2
    eY1, eY2 = float(f32(0.1)), float(f32(0.3))
    hand = (eY1 * (1 - 0.5) * 0.37 + eY2 * (1 - 0.25) * 0.81) / dxdy
    assert abs((FlR.sum() + FlT.sum()) / hand - 1) < 1e-12
    # top/bottom split of each voxel's photons (one voxel at a time)
    for v, (eY, af) in ((0, (eY1, 0.5)), (1, (eY2, 0.25))):
        F1 = np.zeros(3)
        F1[v] = 1.0
        x, a = M.init_rng(L, seed=9 + v)
        R, T, _ = O.fluor_reference(sim, F1, x, a, nv)
        tot = eY * (1 - af) / dxdy
        assert abs((R.sum() + T.sum()) / tot - 1) < 1e-12
        assert abs(R.sum() / tot - 0.5) < 4 * 0.5 / math.sqrt(nv)
