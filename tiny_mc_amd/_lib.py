"""ctypes binding of libtiny_mc_gpu.so (include/tiny_mc_gpu.h).

The library is loaded from this package directory, where ``make -C
tiny_mc_amd`` (or ``__graft_entry__.build()``) puts it.  There is no CPU
fallback: every entry point that needs the GPU raises ``TmcError`` when the
library is missing or the call fails.
"""
import ctypes
import os
from dataclasses import dataclass

import numpy as np

HEAT_FRAC_BITS = 33
HEAT2_FRAC_BITS = 43
MAX_SHELLS = 8192
SAMPLER_ANALYTIC = 0
SAMPLER_REJECTION = 1
PRECISION_FP32 = 0
PRECISION_FP32_WEIGHT = 1
PRECISION_FP32_COMPENSATED = 2
PRECISION_FP64 = 3
MATH_HARDWARE = 0
MATH_OWN = 1

_HERE = os.path.dirname(os.path.abspath(__file__))
lib_path = os.path.join(_HERE, "libtiny_mc_gpu.so")

_STATUS = {
    0: "ok",
    -1: "invalid argument",
    -2: "no usable HIP device",
    -3: "HIP runtime error",
    -4: "out of memory",
    -5: "photon index range exceeds 2^63",
    -6: "unsupported option",
    -7: "more shells than the no-allocation scratch holds",
    -8: "request cancelled",
    -9: "request deadline passed",
    -10: "server close timed out",
}
E_CANCELLED = -8
E_DEADLINE = -9
E_TIMEOUT = -10
NOALLOC_MAX_SHELLS = 4096


class TmcError(RuntimeError):
    """A tiny_mc_gpu call failed; ``status`` holds the tmc_status code."""

    def __init__(self, status, what):
        super().__init__(f"{what}: {_STATUS.get(status, 'unknown status')} ({status})")
        self.status = status


class _Params(ctypes.Structure):
    _fields_ = [
        ("mu_a", ctypes.c_float),
        ("mu_s", ctypes.c_float),
        ("shells", ctypes.c_uint32),
        ("microns_per_shell", ctypes.c_float),
        ("photons", ctypes.c_uint64),
    ]


class _Options(ctypes.Structure):
    _fields_ = [
        ("sampler", ctypes.c_uint32),
        ("block", ctypes.c_uint32),
        ("grid", ctypes.c_uint32),
        ("chunk", ctypes.c_uint32),
        ("first_photon", ctypes.c_uint64),
        ("max_interactions", ctypes.c_uint64),
        ("precision", ctypes.c_uint32),
        ("math", ctypes.c_uint32),
        ("weight_table", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


WEIGHT_TABLE_OFF = 0xFFFFFFFF


class _Stats(ctypes.Structure):
    _fields_ = [
        ("photons", ctypes.c_uint64),
        ("interactions", ctypes.c_uint64),
        ("truncated", ctypes.c_uint64),
        ("kernel_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("grid", ctypes.c_uint32),
        ("block", ctypes.c_uint32),
        ("chunk", ctypes.c_uint32),
        ("sampler", ctypes.c_uint32),
        ("precision", ctypes.c_uint32),
        ("math", ctypes.c_uint32),
        ("weight_table", ctypes.c_uint32),
    ]


@dataclass(frozen=True)
class Params:
    """One simulation: the params.h macros (params.h:5-27) as run-time values."""

    mu_a: float = 2.0
    mu_s: float = 20.0
    shells: int = 101
    microns_per_shell: float = 50.0
    photons: int = 32768

    def _c(self):
        return _Params(self.mu_a, self.mu_s, self.shells, self.microns_per_shell, self.photons)


@dataclass
class RunStats:
    photons: int
    interactions: int
    truncated: int
    kernel_ms: float
    total_ms: float
    grid: int
    block: int
    chunk: int
    sampler: int
    precision: int
    math: int
    weight_table: int = 0


_lib = None


def load():
    """Loads libtiny_mc_gpu.so from the package directory (once)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(lib_path):
        raise TmcError(-6, f"{lib_path} is not built (run make -C tiny_mc_amd)")
    lib = ctypes.CDLL(lib_path)
    P = ctypes.POINTER
    u64p = P(ctypes.c_uint64)
    dp = P(ctypes.c_double)
    lib.tmc_options_default.argtypes = [P(_Options)]
    lib.tmc_options_default.restype = None
    lib.tiny_mc_gpu_run.argtypes = [P(_Params), ctypes.c_uint64, dp, dp]
    lib.tiny_mc_gpu_run_ex.argtypes = [P(_Params), ctypes.c_uint64, P(_Options), u64p, P(_Stats)]
    lib.tiny_mc_gpu_run_noalloc.argtypes = [P(_Params), ctypes.c_uint64, P(_Options), u64p, P(_Stats)]
    lib.tiny_mc_gpu_tally_to_heat.argtypes = [P(_Params), u64p, dp, dp]
    lib.tiny_mc_gpu_init.argtypes = [ctypes.c_int]
    lib.tiny_mc_gpu_shutdown.argtypes = []
    lib.tiny_mc_gpu_shutdown.restype = None
    lib.tiny_mc_gpu_device_info.argtypes = [ctypes.c_char_p, ctypes.c_size_t, P(ctypes.c_int)]
    lib.tiny_mc_gpu_strerror.argtypes = [ctypes.c_int]
    lib.tiny_mc_gpu_strerror.restype = ctypes.c_char_p
    u32p = P(ctypes.c_uint32)
    fp = P(ctypes.c_float)
    lib.tiny_mc_gpu_philox.argtypes = [u32p, u32p, ctypes.c_uint32, u32p]
    lib.tiny_mc_gpu_eval_math.argtypes = [fp, ctypes.c_uint32, fp, fp, fp]
    lib.tiny_mc_gpu_check_sqrt.argtypes = [ctypes.c_uint32, ctypes.c_uint64, u64p, P(ctypes.c_uint32)]
    lib.tiny_mc_gpu_check_short_op.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, u64p,
                                               P(ctypes.c_uint32)]
    lib.tiny_mc_gpu_check_shell_index.argtypes = [P(_Params), ctypes.c_uint32, ctypes.c_uint64, u64p,
                                                  P(ctypes.c_uint32), u64p]
    lib.tiny_mc_gpu_eval_step_math.argtypes = [ctypes.c_uint32, fp, ctypes.c_uint32, fp, fp, fp, fp]
    for name in ("tiny_mc_gpu_run", "tiny_mc_gpu_run_ex", "tiny_mc_gpu_run_noalloc", "tiny_mc_gpu_tally_to_heat", "tiny_mc_gpu_init",
                 "tiny_mc_gpu_device_info", "tiny_mc_gpu_philox", "tiny_mc_gpu_eval_math", "tiny_mc_gpu_check_sqrt",
                 "tiny_mc_gpu_eval_step_math", "tiny_mc_gpu_check_short_op", "tiny_mc_gpu_check_shell_index"):
        getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def _check(rc, what):
    if rc != 0:
        raise TmcError(rc, what)


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def run(params: Params, seed: int):
    """heat, heat2 (float64 arrays of params.shells) in the reference's units."""
    lib = load()
    heat = np.zeros(params.shells, dtype=np.float64)
    heat2 = np.zeros(params.shells, dtype=np.float64)
    c = params._c()
    _check(lib.tiny_mc_gpu_run(ctypes.byref(c), seed, _ptr(heat, ctypes.c_double), _ptr(heat2, ctypes.c_double)),
           "tiny_mc_gpu_run")
    return heat, heat2


def _run_raw(entry, params, seed, sampler, first_photon, max_interactions, block, grid, chunk, precision, math,
             weight_table):
    lib = load()
    tally = np.zeros(4 * params.shells, dtype=np.uint64)
    opt = _Options(sampler, block, grid, chunk, first_photon, max_interactions, precision, math, weight_table, 0)
    st = _Stats()
    c = params._c()
    _check(getattr(lib, entry)(ctypes.byref(c), seed, ctypes.byref(opt), _ptr(tally, ctypes.c_uint64),
                               ctypes.byref(st)), entry)
    stats = RunStats(st.photons, st.interactions, st.truncated, st.kernel_ms, st.total_ms, st.grid, st.block,
                     st.chunk, st.sampler, st.precision, st.math, st.weight_table)
    return tally, stats


def run_raw(params: Params, seed: int, *, sampler=SAMPLER_ANALYTIC, first_photon=0, max_interactions=0,
            block=0, grid=0, chunk=0, precision=PRECISION_FP32, math=MATH_HARDWARE, weight_table=0):
    """Integer tallies (uint64 array of 4*shells: 128-bit (lo, hi) sums) and RunStats."""
    return _run_raw("tiny_mc_gpu_run_ex", params, seed, sampler, first_photon, max_interactions, block, grid, chunk,
                    precision, math, weight_table)


def run_raw_noalloc(params: Params, seed: int, *, sampler=SAMPLER_ANALYTIC, first_photon=0, max_interactions=0,
                    block=0, grid=0, chunk=0, precision=PRECISION_FP32, math=MATH_HARDWARE, weight_table=0):
    """run_raw through tiny_mc_gpu_run_noalloc (static device scratch, no
    allocation; params.shells <= NOALLOC_MAX_SHELLS)."""
    return _run_raw("tiny_mc_gpu_run_noalloc", params, seed, sampler, first_photon, max_interactions, block, grid,
                    chunk, precision, math, weight_table)


def tally_to_heat(params: Params, tally):
    """Host conversion of raw tallies to heat, heat2 (no device work)."""
    lib = load()
    tally = np.ascontiguousarray(tally, dtype=np.uint64)
    heat = np.zeros(params.shells, dtype=np.float64)
    heat2 = np.zeros(params.shells, dtype=np.float64)
    c = params._c()
    _check(lib.tiny_mc_gpu_tally_to_heat(ctypes.byref(c), _ptr(tally, ctypes.c_uint64),
                                         _ptr(heat, ctypes.c_double), _ptr(heat2, ctypes.c_double)),
           "tiny_mc_gpu_tally_to_heat")
    return heat, heat2


def philox(ctr, key):
    """Device Philox4x32-10 on n counters (n x 4 uint32) and keys (n x 2 uint32)."""
    lib = load()
    ctr = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1, 4)
    key = np.ascontiguousarray(key, dtype=np.uint32).reshape(-1, 2)
    assert ctr.shape[0] == key.shape[0]
    out = np.zeros_like(ctr)
    _check(lib.tiny_mc_gpu_philox(_ptr(ctr, ctypes.c_uint32), _ptr(key, ctypes.c_uint32), ctr.shape[0],
                                  _ptr(out, ctypes.c_uint32)), "tiny_mc_gpu_philox")
    return out


def eval_step_math(u, math=MATH_HARDWARE):
    """The walk's step length -ln(u), sin/cos(2 pi u) and the analytic
    sampler's sin(theta) = sqrt(1 - (2u - 1)^2) in math mode `math` on
    inputs u (each ((r >> 8) + 1) 2^-24): float32 arrays (t, sin, cos,
    sin_theta)."""
    lib = load()
    u = np.ascontiguousarray(u, dtype=np.float32)
    t, sn, cs, sp = (np.zeros_like(u) for _ in range(4))
    fp = ctypes.c_float
    _check(lib.tiny_mc_gpu_eval_step_math(math, _ptr(u, fp), u.size, _ptr(t, fp), _ptr(sn, fp), _ptr(cs, fp),
                                          _ptr(sp, fp)), "tiny_mc_gpu_eval_step_math")
    return t, sn, cs, sp


def step_math_tables(math=MATH_HARDWARE):
    """(t, sin, cos, sin_theta) of the walk for every uniform it can draw,
    indexed by r >> 8 (2^24 entries each)."""
    u = ((np.arange(1 << 24, dtype=np.float64) + 1.0) * 2.0 ** -24).astype(np.float32)
    return eval_step_math(u, math)


def check_sqrt(first_bits, count):
    """Mismatches of the walk's short square root against the device IEEE
    sqrt over float bit patterns [first_bits, first_bits + count):
    (count, lowest mismatching pattern)."""
    lib = load()
    n = ctypes.c_uint64()
    f = ctypes.c_uint32()
    _check(lib.tiny_mc_gpu_check_sqrt(first_bits, count, ctypes.byref(n), ctypes.byref(f)), "tiny_mc_gpu_check_sqrt")
    return n.value, f.value


CHECK_SQRT = 0
CHECK_DIV_TENTH_SHORT = 1
CHECK_DIV_TENTH = 2


def check_shell_index(params, first_bits, count):
    """The walk's short shell computation for `params` on squared radii with
    float bit patterns [first_bits, first_bits + count): (violations,
    lowest violating pattern, radii recomputed with the IEEE sqrt)."""
    lib = load()
    n = ctypes.c_uint64()
    f = ctypes.c_uint32()
    u = ctypes.c_uint64()
    p = params._c()
    _check(lib.tiny_mc_gpu_check_shell_index(ctypes.byref(p), first_bits, count, ctypes.byref(n), ctypes.byref(f),
                                             ctypes.byref(u)), "tiny_mc_gpu_check_shell_index")
    return n.value, f.value, u.value


def check_short_op(op, first_bits, count):
    """Mismatches of one of the walk's short forms (CHECK_*) against the
    device IEEE operation over float bit patterns [first_bits, first_bits +
    count): (count, lowest mismatching pattern)."""
    lib = load()
    n = ctypes.c_uint64()
    f = ctypes.c_uint32()
    _check(lib.tiny_mc_gpu_check_short_op(op, first_bits, count, ctypes.byref(n), ctypes.byref(f)),
           "tiny_mc_gpu_check_short_op")
    return n.value, f.value


def eval_math(x):
    """Device log(x), sin(2 pi x), cos(2 pi x) of the walk."""
    lib = load()
    x = np.ascontiguousarray(x, dtype=np.float32)
    lg, sn, cs = (np.zeros_like(x) for _ in range(3))
    fp = ctypes.c_float
    _check(lib.tiny_mc_gpu_eval_math(_ptr(x, fp), x.size, _ptr(lg, fp), _ptr(sn, fp), _ptr(cs, fp)),
           "tiny_mc_gpu_eval_math")
    return lg, sn, cs


def device_info():
    lib = load()
    buf = ctypes.create_string_buffer(128)
    cus = ctypes.c_int()
    _check(lib.tiny_mc_gpu_device_info(buf, len(buf), ctypes.byref(cus)), "tiny_mc_gpu_device_info")
    return buf.value.decode(), cus.value


def shutdown():
    if _lib is not None:
        _lib.tiny_mc_gpu_shutdown()


# ---------------------------------------------------------------------------
# Batched sweep (tmc_batch_*)
# ---------------------------------------------------------------------------


class _Set(ctypes.Structure):
    _fields_ = [("params", _Params), ("seed", ctypes.c_uint64), ("first_photon", ctypes.c_uint64)]


class _BatchOptions(ctypes.Structure):
    _fields_ = [
        ("sampler", ctypes.c_uint32),
        ("block", ctypes.c_uint32),
        ("waves_per_simd", ctypes.c_uint32),
        ("copies", ctypes.c_uint32),
        ("chunk_work", ctypes.c_uint64),
        ("max_interactions", ctypes.c_uint64),
        ("use_graph", ctypes.c_uint32),
        ("math", ctypes.c_uint32),
    ]


class _BatchStats(ctypes.Structure):
    _fields_ = [
        ("sets", ctypes.c_uint64),
        ("photons", ctypes.c_uint64),
        ("interactions", ctypes.c_uint64),
        ("truncated", ctypes.c_uint64),
        ("chunks", ctypes.c_uint64),
        ("chunk_work", ctypes.c_uint64),
        ("expected_interactions", ctypes.c_double),
        ("kernel_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("grid", ctypes.c_uint32),
        ("block", ctypes.c_uint32),
        ("copies", ctypes.c_uint32),
        ("waves_per_simd", ctypes.c_uint32),
        ("idle_lane_steps", ctypes.c_uint64),
        ("passes", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


@dataclass
class BatchStats:
    sets: int
    photons: int
    interactions: int
    truncated: int
    chunks: int
    chunk_work: int
    expected_interactions: float
    kernel_ms: float
    total_ms: float
    grid: int
    block: int
    copies: int
    waves_per_simd: int
    idle_lane_steps: int
    passes: int = 1

    @property
    def lane_utilisation(self):
        """Share of lane-steps that walked a photon while waves stepped."""
        total = self.interactions + self.idle_lane_steps
        return self.interactions / total if total else 0.0


@dataclass
class BatchOptions:
    sampler: int = SAMPLER_ANALYTIC
    block: int = 0
    waves_per_simd: int = 0
    copies: int = 0
    chunk_work: int = 0
    max_interactions: int = 0
    use_graph: bool = False
    math: int = MATH_HARDWARE

    def _c(self):
        return _BatchOptions(self.sampler, self.block, self.waves_per_simd, self.copies, self.chunk_work,
                             self.max_interactions, int(self.use_graph), self.math)


def _bind_batch(lib):
    if getattr(lib, "_tmc_batch_bound", False):
        return
    P = ctypes.POINTER
    lib.tmc_batch_tally_words.argtypes = [P(_Set), ctypes.c_uint32]
    lib.tmc_batch_tally_words.restype = ctypes.c_uint64
    lib.tmc_batch_plan_create.argtypes = [P(_Set), ctypes.c_uint32, P(_BatchOptions), P(ctypes.c_void_p)]
    lib.tmc_batch_plan_create.restype = ctypes.c_int
    lib.tmc_batch_plan_run.argtypes = [ctypes.c_void_p, P(ctypes.c_uint64), P(_BatchStats)]
    lib.tmc_batch_plan_run.restype = ctypes.c_int
    lib.tmc_batch_plan_destroy.argtypes = [ctypes.c_void_p]
    lib.tmc_batch_plan_destroy.restype = None
    lib.tmc_batch_run.argtypes = [P(_Set), ctypes.c_uint32, P(_BatchOptions), P(ctypes.c_uint64), P(_BatchStats)]
    lib.tmc_batch_run.restype = ctypes.c_int
    lib.tmc_batch_autotune.argtypes = [P(_Set), ctypes.c_uint32, P(_BatchOptions), ctypes.c_uint32,
                                       P(_BatchOptions)]
    lib.tmc_batch_autotune.restype = ctypes.c_int
    lib.tmc_batch_options_default.argtypes = [P(_BatchOptions)]
    lib.tmc_batch_options_default.restype = None
    lib._tmc_batch_bound = True


def _sets_array(sets):
    """sets: iterable of (Params, seed) or (Params, seed, first_photon)."""
    sets = list(sets)
    arr = (_Set * max(1, len(sets)))()
    for i, s in enumerate(sets):
        p, seed = s[0], s[1]
        first = s[2] if len(s) > 2 else 0
        arr[i] = _Set(p._c(), seed, first)
    return arr, len(sets)


def _batch_stats(st):
    return BatchStats(st.sets, st.photons, st.interactions, st.truncated, st.chunks, st.chunk_work,
                      st.expected_interactions, st.kernel_ms, st.total_ms, st.grid, st.block, st.copies,
                      st.waves_per_simd, st.idle_lane_steps, st.passes)


def split_tallies(sets, flat):
    """Per-set views (4*shells words each) of concatenated batch tallies."""
    out, off = [], 0
    for s in sets:
        n = 4 * s[0].shells
        out.append(flat[off:off + n])
        off += n
    return out


class BatchPlan:
    """A batch uploaded once and run many times (optionally from a HIP graph)."""

    def __init__(self, sets, options: BatchOptions = None):
        lib = load()
        _bind_batch(lib)
        self._lib = lib
        self.sets = list(sets)
        self._arr, self._n = _sets_array(self.sets)
        self.tally_words = lib.tmc_batch_tally_words(self._arr, self._n)
        opt = (options or BatchOptions())._c()
        h = ctypes.c_void_p()
        _check(lib.tmc_batch_plan_create(self._arr, self._n, ctypes.byref(opt), ctypes.byref(h)),
               "tmc_batch_plan_create")
        self._h = h

    def run(self):
        t = np.zeros(max(1, self.tally_words), dtype=np.uint64)
        st = _BatchStats()
        _check(self._lib.tmc_batch_plan_run(self._h, _ptr(t, ctypes.c_uint64), ctypes.byref(st)),
               "tmc_batch_plan_run")
        return t[:self.tally_words], _batch_stats(st)

    def close(self):
        if self._h:
            self._lib.tmc_batch_plan_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_run(sets, options: BatchOptions = None):
    """Runs a batch once: (concatenated tallies, BatchStats)."""
    lib = load()
    _bind_batch(lib)
    sets = list(sets)
    arr, n = _sets_array(sets)
    words = lib.tmc_batch_tally_words(arr, n)
    t = np.zeros(max(1, words), dtype=np.uint64)
    st = _BatchStats()
    opt = (options or BatchOptions())._c()
    _check(lib.tmc_batch_run(arr, n, ctypes.byref(opt), _ptr(t, ctypes.c_uint64), ctypes.byref(st)),
           "tmc_batch_run")
    return t[:words], _batch_stats(st)


def batch_autotune(sets, base: BatchOptions = None, sample_divisor=16):
    lib = load()
    _bind_batch(lib)
    arr, n = _sets_array(sets)
    b = (base or BatchOptions())._c()
    out = _BatchOptions()
    _check(lib.tmc_batch_autotune(arr, n, ctypes.byref(b), sample_divisor, ctypes.byref(out)),
           "tmc_batch_autotune")
    return BatchOptions(out.sampler, out.block, out.waves_per_simd, out.copies, out.chunk_work,
                        out.max_interactions, bool(out.use_graph), out.math)


def expected_interactions(params: Params):
    """Expected absorptions per photon (host estimate from the weight cadence)."""
    lib = load()
    lib.tiny_mc_gpu_expected_interactions.argtypes = [ctypes.POINTER(_Params)]
    lib.tiny_mc_gpu_expected_interactions.restype = ctypes.c_double
    c = params._c()
    return lib.tiny_mc_gpu_expected_interactions(ctypes.byref(c))


# ---------------------------------------------------------------------------
# Serving (tmc_serve_*)
# ---------------------------------------------------------------------------

PENDING = 1


class _ServeOptions(ctypes.Structure):
    _fields_ = [
        ("sampler", ctypes.c_uint32),
        ("max_shells", ctypes.c_uint32),
        ("max_requests", ctypes.c_uint32),
        ("block", ctypes.c_uint32),
        ("tally_words", ctypes.c_uint64),
        ("ring_chunks", ctypes.c_uint64),
        ("chunk_work", ctypes.c_uint64),
        ("max_photons_per_lane", ctypes.c_uint32),
        ("idle_ms", ctypes.c_uint32),
        ("math", ctypes.c_uint32),
        ("max_chunk_scale", ctypes.c_uint32),
        ("reserve_per_cu", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class _RequestStats(ctypes.Structure):
    _fields_ = [
        ("photons", ctypes.c_uint64),
        ("interactions", ctypes.c_uint64),
        ("submit_s", ctypes.c_double),
        ("admit_s", ctypes.c_double),
        ("done_s", ctypes.c_double),
        ("latency_ms", ctypes.c_double),
    ]


class _ServeStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64),
        ("admitted", ctypes.c_uint64),
        ("completed", ctypes.c_uint64),
        ("photons_completed", ctypes.c_uint64),
        ("interactions_completed", ctypes.c_uint64),
        ("kernel_launches", ctypes.c_uint64),
        ("chunks_published", ctypes.c_uint64),
        ("chunks_claimed", ctypes.c_uint64),
        ("pending", ctypes.c_uint64),
        ("active", ctypes.c_uint64),
        ("elapsed_s", ctypes.c_double),
        ("cancelled", ctypes.c_uint64),
        ("expired", ctypes.c_uint64),
        ("grid", ctypes.c_uint32),
        ("block", ctypes.c_uint32),
    ]


@dataclass
class RequestStats:
    photons: int
    interactions: int
    submit_s: float
    admit_s: float
    done_s: float
    latency_ms: float


def _bind_serve(lib):
    if getattr(lib, "_tmc_serve_bound", False):
        return
    P = ctypes.POINTER
    lib.tmc_serve_open.argtypes = [P(_ServeOptions), P(ctypes.c_void_p)]
    lib.tmc_serve_submit.argtypes = [ctypes.c_void_p, P(_Set), P(ctypes.c_uint64)]
    lib.tmc_serve_submit_ex.argtypes = [ctypes.c_void_p, P(_Set), ctypes.c_double, P(ctypes.c_uint64)]
    lib.tmc_serve_cancel.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    lib.tmc_serve_close_timeout.argtypes = [ctypes.c_void_p, ctypes.c_double]
    lib.tmc_serve_poll.argtypes = [ctypes.c_void_p, ctypes.c_uint64, P(ctypes.c_uint64), P(_RequestStats)]
    lib.tmc_serve_wait.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double, P(ctypes.c_uint64),
                                   P(_RequestStats)]
    lib.tmc_serve_get_stats.argtypes = [ctypes.c_void_p, P(_ServeStats)]
    lib.tmc_serve_close.argtypes = [ctypes.c_void_p]
    for n in ("tmc_serve_open", "tmc_serve_submit", "tmc_serve_submit_ex", "tmc_serve_cancel", "tmc_serve_poll",
              "tmc_serve_wait", "tmc_serve_get_stats", "tmc_serve_close", "tmc_serve_close_timeout"):
        getattr(lib, n).restype = ctypes.c_int
    lib._tmc_serve_bound = True


class Server:
    """Request loop over a resident persistent kernel (tmc_serve_*).

    Used as a context manager it closes with a time limit (``close_timeout``
    seconds for unfinished requests, then cancelled ones drain), so a
    request that never finishes fails the caller instead of holding it."""

    def __init__(self, sampler=SAMPLER_ANALYTIC, max_shells=0, max_requests=0, block=0, tally_words=0,
                 ring_chunks=0, chunk_work=0, max_photons_per_lane=0, idle_ms=0, math=MATH_HARDWARE,
                 max_chunk_scale=0, reserve_per_cu=0, close_timeout=120.0):
        lib = load()
        _bind_serve(lib)
        self._lib = lib
        o = _ServeOptions(sampler, max_shells, max_requests, block, tally_words, ring_chunks, chunk_work,
                          max_photons_per_lane, idle_ms, math, max_chunk_scale, reserve_per_cu, 0)
        h = ctypes.c_void_p()
        _check(lib.tmc_serve_open(ctypes.byref(o), ctypes.byref(h)), "tmc_serve_open")
        self._h = h
        self._shells = {}
        self.close_timeout = close_timeout

    def submit(self, params: Params, seed: int, first_photon: int = 0, deadline=None):
        """Queues a request (deadline: seconds after submission, or None)."""
        s = _Set(params._c(), seed, first_photon)
        rid = ctypes.c_uint64()
        _check(self._lib.tmc_serve_submit_ex(self._h, ctypes.byref(s), float(deadline or 0.0), ctypes.byref(rid)),
               "tmc_serve_submit_ex")
        self._shells[rid.value] = params.shells
        return rid.value

    def cancel(self, rid):
        _check(self._lib.tmc_serve_cancel(self._h, rid), "tmc_serve_cancel")

    def _result(self, rc, rid, t, st, what):
        if rc == PENDING:
            return None
        if rc in (0, E_CANCELLED, E_DEADLINE):  # the outcome is handed over once
            self._shells.pop(rid, None)
        _check(rc, what)
        return t, RequestStats(st.photons, st.interactions, st.submit_s, st.admit_s, st.done_s, st.latency_ms)

    def poll(self, rid):
        """(tally, RequestStats) once done, None while pending; TmcError with
        status E_CANCELLED / E_DEADLINE for a cancelled request."""
        t = np.zeros(4 * self._shells[rid], dtype=np.uint64)
        st = _RequestStats()
        rc = self._lib.tmc_serve_poll(self._h, rid, _ptr(t, ctypes.c_uint64), ctypes.byref(st))
        return self._result(rc, rid, t, st, "tmc_serve_poll")

    def wait(self, rid, timeout=-1.0):
        """As poll, blocking up to `timeout` seconds (< 0: no limit)."""
        t = np.zeros(4 * self._shells[rid], dtype=np.uint64)
        st = _RequestStats()
        rc = self._lib.tmc_serve_wait(self._h, rid, timeout, _ptr(t, ctypes.c_uint64), ctypes.byref(st))
        return self._result(rc, rid, t, st, "tmc_serve_wait")

    def stats(self):
        st = _ServeStats()
        _check(self._lib.tmc_serve_get_stats(self._h, ctypes.byref(st)), "tmc_serve_get_stats")
        return {f: getattr(st, f) for f, _ in _ServeStats._fields_}

    def close(self, timeout=None):
        """Finishes the submitted requests and frees the server; with a
        timeout, cancels what is unfinished after `timeout` seconds and
        raises TmcError(E_TIMEOUT) if the kernel has not left within as long
        again (the server then stays open; close may be called again)."""
        if self._h:
            rc = self._lib.tmc_serve_close_timeout(self._h, -1.0 if timeout is None else float(timeout))
            if rc == E_TIMEOUT:
                _check(rc, "tmc_serve_close_timeout")
            self._h = None
            _check(rc, "tmc_serve_close_timeout")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close(self.close_timeout)
