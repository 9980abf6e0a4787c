"""tiny_mc_amd: the MI355X photon-transport path of tiny_mc.

The product is the C ABI in ``include/tiny_mc_gpu.h`` (``libtiny_mc_gpu.so``,
HIP kernels for gfx950).  This package is its Python host binding, used by
the tests and the benchmark, and mirrors the reference's interface: the
params.h macros become ``Params`` fields and the photon() loop of
tiny_mc.c:43-49 becomes one ``run`` call.
"""
from ._lib import (  # noqa: F401
    PENDING,
    E_CANCELLED,
    E_DEADLINE,
    E_TIMEOUT,
    BatchOptions,
    RequestStats,
    Server,
    expected_interactions,
    BatchPlan,
    BatchStats,
    batch_autotune,
    batch_run,
    split_tallies,
    HEAT2_FRAC_BITS,
    HEAT_FRAC_BITS,
    MATH_HARDWARE,
    MATH_OWN,
    MAX_SHELLS,
    PRECISION_FP32,
    PRECISION_FP32_COMPENSATED,
    PRECISION_FP32_WEIGHT,
    PRECISION_FP64,
    SAMPLER_ANALYTIC,
    SAMPLER_REJECTION,
    Params,
    RunStats,
    TmcError,
    check_sqrt,
    check_short_op,
    check_shell_index,
    CHECK_SQRT,
    CHECK_DIV_TENTH_SHORT,
    CHECK_DIV_TENTH,
    device_info,
    eval_math,
    eval_step_math,
    step_math_tables,
    lib_path,
    load,
    philox,
    run,
    run_raw,
    run_raw_noalloc,
    WEIGHT_TABLE_OFF,
    NOALLOC_MAX_SHELLS,
    shutdown,
    tally_to_heat,
)
from . import serve_load  # noqa: F401
from .sweep import sweep_sets  # noqa: F401
from .table import format_table, table_rows  # noqa: F401
