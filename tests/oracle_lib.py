"""ctypes access to the oracle (oracle/liboracle.so) for the tests.

The oracle is the checker: CPU restatement of the reference (glibc rand()),
the FP64 walk on the Philox stream, and the FP32 mirror of the GPU kernel.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

SAMPLER_ANALYTIC = 0
SAMPLER_REJECTION = 1


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("mu_a", ctypes.c_float),
        ("mu_s", ctypes.c_float),
        ("shells", ctypes.c_uint32),
        ("microns_per_shell", ctypes.c_float),
    ]


def threads():
    n = os.cpu_count() or 1
    return max(1, min(16, n))


class Oracle:
    def __init__(self, path=LIB):
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built (make -C oracle)")
        L = ctypes.CDLL(path)
        P = ctypes.POINTER
        L.orc_restate.argtypes = [P(OrcParams), ctypes.c_uint, ctypes.c_uint32, P(ctypes.c_float),
                                  P(ctypes.c_float), P(ctypes.c_uint64)]
        L.orc_restate.restype = ctypes.c_int
        L.orc_restate_table.argtypes = [P(OrcParams), ctypes.c_uint, ctypes.c_uint32, ctypes.c_char_p,
                                        ctypes.c_size_t]
        L.orc_restate_table.restype = ctypes.c_long
        L.orc_walk64.argtypes = [P(OrcParams), ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                 ctypes.c_uint64, ctypes.c_int, P(ctypes.c_double), P(ctypes.c_double),
                                 P(ctypes.c_uint64)]
        L.orc_walk64.restype = ctypes.c_int
        L.orc_mirror32.argtypes = [P(OrcParams), ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                   ctypes.c_uint64, ctypes.c_int, P(ctypes.c_uint64), P(ctypes.c_uint64)]
        L.orc_mirror32.restype = ctypes.c_int
        fp = P(ctypes.c_float)
        L.orc_mirror32_tab.argtypes = [P(OrcParams), ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                       ctypes.c_uint64, ctypes.c_int, fp, fp, fp, fp, P(ctypes.c_uint64),
                                       P(ctypes.c_uint64)]
        L.orc_mirror32_tab.restype = ctypes.c_int
        L.orc_philox.argtypes = [P(ctypes.c_uint32), P(ctypes.c_uint32), P(ctypes.c_uint32)]
        L.orc_philox.restype = None
        L.orc_u01_f.argtypes = [ctypes.c_uint32]
        L.orc_u01_f.restype = ctypes.c_float
        L.orc_log32.argtypes = [ctypes.c_float]
        L.orc_log32.restype = ctypes.c_float
        L.orc_sincos2pi32.argtypes = [ctypes.c_float, P(ctypes.c_float), P(ctypes.c_float)]
        L.orc_sincos2pi32.restype = None
        L.orc_derived32.argtypes = [P(OrcParams), P(ctypes.c_float), P(ctypes.c_float)]
        L.orc_derived32.restype = None
        L.orc_trace64.argtypes = [P(OrcParams), ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64,
                                  P(ctypes.c_uint32), P(ctypes.c_double)]
        L.orc_trace64.restype = ctypes.c_uint64
        L.orc_directions64.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64,
                                       P(ctypes.c_double)]
        L.orc_directions64.restype = ctypes.c_int
        L.orc_math32_errors.argtypes = [P(ctypes.c_double)] * 3
        L.orc_math32_errors.restype = None
        L.orc_eval_math32.argtypes = [P(ctypes.c_float), ctypes.c_uint64, P(ctypes.c_float), P(ctypes.c_float),
                                      P(ctypes.c_float)]
        L.orc_eval_math32.restype = None
        self.L = L

    @staticmethod
    def params(mu_a=2.0, mu_s=20.0, shells=101, microns_per_shell=50.0):
        return OrcParams(mu_a, mu_s, shells, microns_per_shell)

    def restate(self, p, seed, photons):
        h = np.zeros(p.shells, np.float32)
        h2 = np.zeros(p.shells, np.float32)
        n = ctypes.c_uint64()
        rc = self.L.orc_restate(ctypes.byref(p), seed, photons, h.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                h2.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(n))
        assert rc == 0
        return h, h2, n.value

    def restate_table(self, p, seed, photons):
        need = self.L.orc_restate_table(ctypes.byref(p), seed, photons, None, 0)
        assert need > 0
        buf = ctypes.create_string_buffer(need + 1)
        n = self.L.orc_restate_table(ctypes.byref(p), seed, photons, buf, need + 1)
        assert n == need
        return buf.value.decode()

    def walk64(self, p, seed, count, first=0, sampler=SAMPLER_ANALYTIC, max_k=0, nthreads=None):
        h = np.zeros(p.shells)
        h2 = np.zeros(p.shells)
        n = ctypes.c_uint64()
        dp = ctypes.POINTER(ctypes.c_double)
        rc = self.L.orc_walk64(ctypes.byref(p), seed, first, count, sampler, max_k, nthreads or threads(),
                               h.ctypes.data_as(dp), h2.ctypes.data_as(dp), ctypes.byref(n))
        assert rc == 0
        return h, h2, n.value

    def mirror32(self, p, seed, count, first=0, sampler=SAMPLER_ANALYTIC, max_k=0, nthreads=None, tables=None):
        """FP32 mirror; tables = (t, sin, cos, sin_theta) float32 arrays of
        2^24 entries indexed by r >> 8 for the kernel's hardware-math mode,
        None for the own polynomials."""
        t = np.zeros(4 * p.shells, np.uint64)
        n = ctypes.c_uint64()
        u64p = ctypes.POINTER(ctypes.c_uint64)
        if tables is None:
            rc = self.L.orc_mirror32(ctypes.byref(p), seed, first, count, sampler, max_k, nthreads or threads(),
                                     t.ctypes.data_as(u64p), ctypes.byref(n))
        else:
            tt, ts, tc, tp = (np.ascontiguousarray(a, dtype=np.float32) for a in tables)
            assert tt.size == ts.size == tc.size == tp.size == 1 << 24
            fp = ctypes.POINTER(ctypes.c_float)
            rc = self.L.orc_mirror32_tab(ctypes.byref(p), seed, first, count, sampler, max_k, nthreads or threads(),
                                         tt.ctypes.data_as(fp), ts.ctypes.data_as(fp), tc.ctypes.data_as(fp),
                                         tp.ctypes.data_as(fp), t.ctypes.data_as(u64p), ctypes.byref(n))
        assert rc == 0
        return t, n.value

    def philox(self, ctr, key):
        c = (ctypes.c_uint32 * 4)(*ctr)
        k = (ctypes.c_uint32 * 2)(*key)
        o = (ctypes.c_uint32 * 4)()
        self.L.orc_philox(c, k, o)
        return list(o)

    def derived32(self, p):
        a, s = ctypes.c_float(), ctypes.c_float()
        self.L.orc_derived32(ctypes.byref(p), ctypes.byref(a), ctypes.byref(s))
        return a.value, s.value

    def trace64(self, p, seed, photon, sampler=SAMPLER_ANALYTIC, cap=1 << 16):
        sh = np.zeros(cap, np.uint32)
        w = np.zeros(cap)
        n = self.L.orc_trace64(ctypes.byref(p), seed, photon, sampler, cap,
                               sh.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                               w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        m = min(n, cap)
        return sh[:m], w[:m], n

    def directions64(self, seed, photon, sampler, n):
        out = np.zeros(3 * n)
        self.L.orc_directions64(seed, photon, sampler, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out.reshape(n, 3)

    def eval_math32(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        lg, sn, cs = (np.zeros_like(x) for _ in range(3))
        fp = ctypes.POINTER(ctypes.c_float)
        self.L.orc_eval_math32(x.ctypes.data_as(fp), x.size, lg.ctypes.data_as(fp), sn.ctypes.data_as(fp),
                               cs.ctypes.data_as(fp))
        return lg, sn, cs

    def math32_errors(self):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self.L.orc_math32_errors(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value


_orc = None


def load():
    global _orc
    if _orc is None:
        _orc = Oracle()
    return _orc


def tally_heat(tally, shells, mu_a, mu_s):
    """128-bit raw tallies -> heat, heat2 in the reference's units (float64),
    computed independently of the product library."""
    t = np.asarray(tally, dtype=np.uint64)
    S = shells
    # the parameters are binary32 in the C ABI; the conversion is in double
    mu_a, mu_s = float(np.float32(mu_a)), float(np.float32(mu_s))
    albedo = mu_s / (mu_s + mu_a)
    oma = 1.0 - albedo

    def u128(lo, hi):
        return hi.astype(np.float64) * 2.0 ** 64 + lo.astype(np.float64)

    w = u128(t[0:2 * S:2], t[1:2 * S:2])
    w2 = u128(t[2 * S::2], t[2 * S + 1::2])
    return oma * w * 2.0 ** -33, oma * oma * w2 * 2.0 ** -43


# roulette survival: (r >> 8) < 0x199999, i.e. u = ((r >> 8) + 1) 2^-24 <= 0.1f
P_SURVIVE = 1677721 / 2**24


def cadence_moments(mu_a, mu_s, cycles=40):
    """Exact mean and variance, per photon, of the absorptions, of the
    absorbed energy sum (1 - a) w and of the sum of fixed-point heat2 images
    RN(w w) 2^43, from the binary32 weight sequence: cycle 0 starts at
    w = 1; each absorption deposits w, then w = RN(w a); once w < 0.001 the
    photon survives with probability P_SURVIVE and the next cycle starts at
    RN(w / 0.1f).  G survivals (P = p^G (1 - p)) walk cycles 0..G."""
    f = np.float32
    a32 = f(f(mu_s) / f(f(mu_s) + f(mu_a)))
    oma = 1.0 - mu_s / (mu_s + mu_a)
    w = f(1.0)
    lens, energy, sq = [], [], []
    for _ in range(cycles):
        n, e, q = 0, 0.0, 0
        while True:
            n += 1
            e += float(w)
            q += int(float(f(w * w)) * 2.0**43)  # exact: w*w >= 2^-20 here
            w = f(w * a32)
            if w < f(0.001):
                break
        lens.append(n)
        energy.append(oma * e)
        sq.append(q)
        w = f(w / f(0.1))
    g = np.arange(cycles)
    pg = P_SURVIVE**g * (1.0 - P_SURVIVE)
    out = []
    for per_cycle in (lens, energy, sq):
        tot = np.cumsum(np.array(per_cycle, dtype=np.float64))
        m = float((pg * tot).sum())
        out.append((m, float((pg * tot * tot).sum()) - m * m))
    return out
