"""Host-side mirror of the reference's assembly entry point.

``assemble_gpu(operator, test_functions, trial_functions, config)`` follows
assemble_gpu (assamble_gpu.jl:25-137): it assembles the dense Galerkin matrix
of ``operator`` on the device and returns it as a host column-major complex
matrix, the same value BEAST's ``assemble(op, X, Y)`` returns.  Everything
runs through libbmi.so; if the library or the GPU is missing it raises.

``Assembler`` keeps mesh, spaces, plan and matrix resident on the device for
repeated assemblies (benchmarks, serving).
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from . import _bmi
from .operators import HH3D_SINGLELAYER, MW_DOUBLELAYER, MW_SINGLELAYER, Operator
from .spaces import P1, RWG, Space


@dataclasses.dataclass
class QuadratureRules:
    """Quadrature tables used in place of the rule orders of a configuration
    (bmi_rules): the far-pair triangle rules of test and trial cells -- chart
    coordinates (u, v) of the point v3 + u (v1 - v3) + v (v2 - v3) and
    weights on the reference triangle (area 1/2), at most 4 points each --
    and the Gauss-Legendre rules on [0, 1] of the Sauter-Schwab hypercube for
    common vertex, edge and face (at most 8 points each).  BEAST's quaddata
    holds the far rules per cell, as Cartesian points and weights times the
    Jacobian 2A (GPU_scheduler.jl:61-90); ``reference_rule`` recovers the
    reference rule from one cell's points."""

    test_uv: np.ndarray
    test_weights: np.ndarray
    trial_uv: np.ndarray
    trial_weights: np.ndarray
    gauss_legendre: tuple  # ((nodes, weights) for CV, CE, CF)

    def __post_init__(self):
        self.test_uv = np.asarray(self.test_uv, dtype=np.float64).reshape(-1, 2)
        self.trial_uv = np.asarray(self.trial_uv, dtype=np.float64).reshape(-1, 2)
        self.test_weights = np.asarray(self.test_weights, dtype=np.float64).ravel()
        self.trial_weights = np.asarray(self.trial_weights, dtype=np.float64).ravel()
        if self.test_uv.shape[0] != self.test_weights.size or self.trial_uv.shape[0] != self.trial_weights.size:
            raise ValueError("one weight per rule point")
        if len(self.gauss_legendre) != 3:
            raise ValueError("three Gauss-Legendre rules: common vertex, edge, face")
        self.gauss_legendre = tuple((np.asarray(x, dtype=np.float64).ravel(), np.asarray(w, dtype=np.float64).ravel())
                                    for x, w in self.gauss_legendre)

    @staticmethod
    def reference_rule(vertices, points, weights):
        """(uv, reference weights) of a rule given on one cell (vertices 3 x 3,
        rows v1 v2 v3) by Cartesian points and weights that include the
        Jacobian 2A, as BEAST's quaddata stores it."""
        v = np.asarray(vertices, dtype=np.float64)
        t = np.stack([v[0] - v[2], v[1] - v[2]], axis=1)  # 3 x 2
        uv, *_ = np.linalg.lstsq(t, (np.asarray(points, dtype=np.float64) - v[2]).T, rcond=None)
        two_area = np.linalg.norm(np.cross(t[:, 0], t[:, 1]))
        return uv.T, np.asarray(weights, dtype=np.float64) / two_area

    def c(self) -> _bmi.Rules:
        r = _bmi.Rules()
        r.test_points, r.trial_points = self.test_uv.shape[0], self.trial_uv.shape[0]
        if not (1 <= r.test_points <= 4 and 1 <= r.trial_points <= 4):
            raise ValueError("1..4 far-rule points per cell")
        for i in range(r.test_points):
            r.test_uv[i][0], r.test_uv[i][1], r.test_weight[i] = self.test_uv[i, 0], self.test_uv[i, 1], self.test_weights[i]
        for i in range(r.trial_points):
            r.trial_uv[i][0], r.trial_uv[i][1], r.trial_weight[i] = (self.trial_uv[i, 0], self.trial_uv[i, 1],
                                                                     self.trial_weights[i])
        for g, (x, w) in enumerate(self.gauss_legendre):
            if not 1 <= x.size <= 8 or x.size != w.size:
                raise ValueError("1..8 Gauss-Legendre points, one weight each")
            r.gl_points[g] = x.size
            for i in range(x.size):
                r.gl_node[g][i], r.gl_weight[g][i] = x[i], w[i]
        return r


@dataclasses.dataclass
class GPUConfiguration:
    """Options of one assembly; mirrors the reference's GPUConfiguration
    (utils/configuration.jl:11-26) for the fields that change what is computed
    or where the result lives.

    writeBackStrategy: "gpu" assembles the whole matrix on the device and
        copies it to the host at the end (GpuWriteBackTrueInstance); "cpu"
        streams it into host memory (GpuWriteBackFalseInstance): row blocks
        of test panels are assembled into device staging blocks of at most
        GPU_budget_pipeline_result bytes together and copied out while the
        next block is assembled, so the matrix need not fit in device memory.
        Both give the same values bitwise.
    GPU_budget_pipeline_result: device bytes of the staging blocks of the "cpu"
        strategy (the reference's field of that name, 24 GiB by default).
    floatType / precision: "fp64" (plain) or "mixed" (FP32 far pairs).
    output: "c64" (ComplexF64) or "c32" (ComplexF32) matrix.
    tile: (panel rows, patch columns) of the far-pair kernels' tile, at most
        (64, 40) on RWG and (128, 20) on Lagrange spaces; 0 takes that
        maximum, the fastest shape measured on every workload.  "auto": the
        shape the tile autotuner (``tune_tile``) times fastest for this
        problem's signature.  Other shapes give the same matrix up to
        summation order.
    """

    writeBackStrategy: str = "gpu"
    GPU_budget_pipeline_result: int = 24 << 30
    precision: str = "fp64"
    output: str = "c64"
    device: int = 0
    ss_order: tuple = (5, 5, 5)
    far_test_order: int = 2
    far_trial_order: int = 3
    near_order: int = 4
    near_factor: float = 2.0
    rules: QuadratureRules | None = None  # replaces far_test_order, far_trial_order and ss_order
    tile: tuple | str = (0, 0)

    def options(self) -> _bmi.Options:
        o = _bmi.default_options()
        o.far_test_order = int(self.far_test_order)
        o.far_trial_order = int(self.far_trial_order)
        for i in range(3):
            o.ss_order[i] = int(self.ss_order[i])
        o.near_order = int(self.near_order)
        if self.precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        o.precision = _bmi.PREC_MIXED if self.precision == "mixed" else _bmi.PREC_FP64
        o.near_factor = float(self.near_factor)
        return o

    def dtype(self) -> int:
        if self.output not in ("c64", "c32"):
            raise ValueError("output must be 'c64' or 'c32'")
        return _bmi.C32 if self.output == "c32" else _bmi.C64


class _Handle:
    _free = None

    def __init__(self):
        self.h = ctypes.c_void_p()

    def close(self):
        if self.h and self.h.value:
            getattr(_bmi.lib(), self._free)(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context(_Handle):
    _free = "bmi_context_destroy"

    def __init__(self, device: int = 0):
        super().__init__()
        _bmi.check(_bmi.lib().bmi_context_create(int(device), ctypes.byref(self.h)))
        self.device = device


class DeviceMesh(_Handle):
    _free = "bmi_mesh_free"

    def __init__(self, ctx: Context, mesh):
        super().__init__()
        self.ctx = ctx
        self._xyz = np.ascontiguousarray(mesh.vertices, dtype=np.float64)
        self._tri = np.ascontiguousarray(mesh.faces, dtype=np.int32)
        _bmi.check(_bmi.lib().bmi_mesh_create(ctx.h, self._xyz.ctypes.data, self._xyz.shape[0],
                                              self._tri.ctypes.data, self._tri.shape[0], ctypes.byref(self.h)))


class DeviceSpace(_Handle):
    _free = "bmi_space_free"

    def __init__(self, dmesh: DeviceMesh, space: Space, tile: tuple = (0, 0)):
        super().__init__()
        self.dmesh = dmesh
        self.space = space
        kind = {P1: _bmi.SPACE_LAGRANGE_C0D1, RWG: _bmi.SPACE_RWG}[space.kind]
        idx = np.ascontiguousarray(space.idx, dtype=np.int64)
        cf = np.ascontiguousarray(space.coeff, dtype=np.float64)
        rows, cols = (int(v) for v in tile)
        _bmi.check(_bmi.lib().bmi_space_create_shaped(dmesh.h, kind, idx.ctypes.data, cf.ctypes.data,
                                                      int(space.nfun), rows, cols, ctypes.byref(self.h)))


def _c_operator(op: Operator) -> _bmi.Operator:
    o = _bmi.Operator()
    o.kind = {HH3D_SINGLELAYER: _bmi.OP_HH3D_SINGLELAYER, MW_SINGLELAYER: _bmi.OP_MW_SINGLELAYER,
              MW_DOUBLELAYER: _bmi.OP_MW_DOUBLELAYER}[op.kind]
    for name in ("alpha", "beta", "gamma"):
        z = complex(getattr(op, name))
        getattr(o, name)[0] = z.real
        getattr(o, name)[1] = z.imag
    return o


class Plan(_Handle):
    _free = "bmi_plan_free"

    def __init__(self, ctx: Context, op: Operator, dX: DeviceSpace, dY: DeviceSpace, config: GPUConfiguration):
        super().__init__()
        self.ctx, self.dX, self.dY = ctx, dX, dY
        o = config.options()
        cop = _c_operator(op)
        if config.rules is None:
            _bmi.check(_bmi.lib().bmi_plan_create(ctx.h, ctypes.byref(cop), dX.h, dY.h, ctypes.byref(o),
                                                  ctypes.byref(self.h)))
        else:
            r = config.rules.c()
            _bmi.check(_bmi.lib().bmi_plan_create_rules(ctx.h, ctypes.byref(cop), dX.h, dY.h, ctypes.byref(o),
                                                        ctypes.byref(r), ctypes.byref(self.h)))

    def stats(self) -> dict:
        s = _bmi.Stats()
        _bmi.check(_bmi.lib().bmi_plan_stats(self.h, ctypes.byref(s)))
        return s.as_dict()


class DeviceMatrix(_Handle):
    _free = "bmi_matrix_free"

    def __init__(self, ctx: Context, rows: int, cols: int, dtype: int = _bmi.C64):
        super().__init__()
        self.ctx, self.rows, self.cols, self.dtype = ctx, int(rows), int(cols), dtype
        _bmi.check(_bmi.lib().bmi_matrix_create(ctx.h, self.rows, self.cols, int(dtype), ctypes.byref(self.h)))

    def device_ptr(self) -> int:
        p = ctypes.c_void_p()
        ld = ctypes.c_int64()
        _bmi.check(_bmi.lib().bmi_matrix_device_ptr(self.h, ctypes.byref(p), ctypes.byref(ld)))
        return p.value

    def apply(self, x: np.ndarray) -> np.ndarray:
        """y = Z x on the device; x is (cols,) or (cols, nrhs) with nrhs <= 8."""
        x = np.asarray(x, dtype=np.complex128)
        vec = x.ndim == 1
        xf = np.asfortranarray(x.reshape(self.cols, -1))
        nrhs = xf.shape[1]
        y = np.empty((self.rows, nrhs), dtype=np.complex128, order="F")
        _bmi.check(_bmi.lib().bmi_matrix_apply(self.h, xf.ctypes.data, y.ctypes.data, nrhs))
        return y[:, 0] if vec else y

    def compare(self, ref: "DeviceMatrix") -> dict:
        """||self - ref||_F, ||ref||_F, max |self - ref| computed on the device."""
        out = (ctypes.c_double * 3)()
        _bmi.check(_bmi.lib().bmi_matrix_compare(self.h, ref.h, out))
        return {"diff_fro": out[0], "ref_fro": out[1], "max_abs": out[2],
                "rel_fro": out[0] / out[1] if out[1] > 0 else float("nan")}

    def to_host(self, dtype: int | None = None) -> np.ndarray:
        dtype = self.dtype if dtype is None else dtype
        out = np.empty((self.rows, self.cols), dtype=np.complex128 if dtype == _bmi.C64 else np.complex64, order="F")
        _bmi.check(_bmi.lib().bmi_matrix_to_host(self.h, out.ctypes.data, int(dtype)))
        return out


# Far-pair tile shapes the autotuner times, per space kind ((0, 0): the
# largest tile, the default)
TILE_CANDIDATES = {RWG: ((0, 0), (48, 40), (64, 28), (32, 40)), P1: ((0, 0), (96, 20), (64, 20), (128, 12))}
_TILE_CACHE: dict = {}


def tile_signature(ctx: Context, op: Operator, X: Space, Y: Space, config: GPUConfiguration) -> tuple:
    """What the fastest tile shape depends on: device, operator kind and
    lossiness, precision and output type, the far rules (orders, or the
    point counts of caller tables), space kind, the test and trial meshes'
    size classes (powers of two of their cell counts) and whether test and
    trial space are one."""
    r = config.rules
    rule = (("tables", r.test_weights.size, r.trial_weights.size) if r is not None
            else (config.far_test_order, config.far_trial_order))
    return (ctx.device, op.kind, complex(op.gamma).real != 0.0, config.precision, config.output, rule, X.kind,
            int(X.mesh.nt).bit_length(), int(Y.mesh.nt).bit_length(), X is Y)


def tune_tile(ctx: Context, op: Operator, X: Space, Y: Space, config: GPUConfiguration,
              matrix: "DeviceMatrix", dmx: DeviceMesh, dmy: DeviceMesh, margin: float = 0.03,
              budget_ms: float = 4.0) -> tuple:
    """The far-pair tile shape with the shortest assembly among
    TILE_CANDIDATES[X.kind], and every candidate's time (ms): each candidate
    builds its spaces and plan on the problem itself and assembles into
    ``matrix`` once to warm up, then is timed by the library's events over
    enough assemblies to fill ``budget_ms`` (2 to 20; the shortest kept).
    Another shape replaces the default (0, 0) only when it is faster by
    more than ``margin`` (a fraction), so timing noise on small problems
    does not flip the choice.  Cached per tile_signature for the process,
    so problems of one signature are timed once.  No counterpart in the
    reference, whose analogous knobs are swept by hand
    (testing/determin_amount_of_threads.jl:44-75)."""
    key = tile_signature(ctx, op, X, Y, config)
    hit = _TILE_CACHE.get(key)
    if hit is not None:
        return hit
    times = {}
    for tile in TILE_CANDIDATES[X.kind]:
        made = []
        try:
            dX = DeviceSpace(dmx, X, tile)
            made.append(dX)
            dY = dX
            if Y is not X:
                dY = DeviceSpace(dmy, Y, tile)
                made.append(dY)
            plan = Plan(ctx, op, dX, dY, config)
            made.append(plan)
            _bmi.check(_bmi.lib().bmi_assemble(plan.h, matrix.h))
            first = plan.stats()["total_ms"]
            best = float("inf")
            for _ in range(min(20, max(2, int(budget_ms / max(first, 1e-3))))):
                _bmi.check(_bmi.lib().bmi_assemble(plan.h, matrix.h))
                best = min(best, plan.stats()["total_ms"])
            times[tile] = best
        finally:
            for o in reversed(made):
                o.close()
    choice = min(times, key=times.get)
    if times[choice] >= times[(0, 0)] * (1.0 - margin):
        choice = (0, 0)
    _TILE_CACHE[key] = (choice, times)
    return choice, times


class Assembler:
    """Resident assembly of one (operator, test space, trial space).  With
    ``resident=False`` no device matrix is allocated: only assemble_to_host
    is available (matrices larger than device memory), and ``tile="auto"``
    takes the default shape (the tuner needs a resident matrix)."""

    def __init__(self, op: Operator, X: Space, Y: Space, config: GPUConfiguration | None = None,
                 ctx: Context | None = None, resident: bool = True):
        self.config = config or GPUConfiguration()
        self.matrix = self.plan = self.dY = self.dX = self.dmy = self.dmx = None
        tile = self.config.tile
        if isinstance(tile, str) and tile != "auto":
            raise ValueError('tile must be (panel rows, patch columns) or "auto"')
        self.ctx = ctx or Context(self.config.device)
        self.tile_times = None
        try:
            self.dmx = DeviceMesh(self.ctx, X.mesh)
            self.dmy = self.dmx if Y.mesh is X.mesh else DeviceMesh(self.ctx, Y.mesh)
            self.shape = (X.nfun, Y.nfun)
            self.matrix = DeviceMatrix(self.ctx, X.nfun, Y.nfun, self.config.dtype()) if resident else None
            if isinstance(tile, str):
                tile = (0, 0)
                if self.matrix is not None:
                    tile, self.tile_times = tune_tile(self.ctx, op, X, Y, self.config, self.matrix, self.dmx,
                                                      self.dmy)
            self.tile = tuple(tile)
            self.dX = DeviceSpace(self.dmx, X, self.tile)
            self.dY = self.dX if Y is X else DeviceSpace(self.dmy, Y, self.tile)
            self.plan = Plan(self.ctx, op, self.dX, self.dY, self.config)
        except BaseException:
            # release what was built, children first (the exception's
            # traceback would otherwise keep it alive)
            self.close()
            raise

    def assemble(self) -> DeviceMatrix:
        if self.matrix is None:
            raise ValueError("this assembler keeps no device matrix: use assemble_to_host")
        _bmi.check(_bmi.lib().bmi_assemble(self.plan.h, self.matrix.h))
        return self.matrix

    def assemble_to_host(self, budget: int | None = None, out: np.ndarray | None = None) -> np.ndarray:
        """Assemble straight into a host column-major matrix through device
        staging blocks of at most ``budget`` bytes together (default: the
        configuration's GPU_budget_pipeline_result)."""
        budget = int(self.config.GPU_budget_pipeline_result if budget is None else budget)
        dt = self.config.dtype()
        want = np.complex128 if dt == _bmi.C64 else np.complex64
        if out is None:
            out = np.empty(self.shape, dtype=want, order="F")
        elif out.shape != self.shape or out.dtype != want or not out.flags.f_contiguous:
            raise ValueError(f"out must be a Fortran-ordered {self.shape} {np.dtype(want).name} array")
        _bmi.check(_bmi.lib().bmi_assemble_to_host(self.plan.h, out.ctypes.data, dt, budget))
        return out

    def stats(self) -> dict:
        return self.plan.stats()

    def close(self):
        for o in (self.matrix, self.plan, self.dY, self.dX, self.dmy, self.dmx):
            if o is not None:
                o.close()


class Batch(_Handle):
    """Many resident assemblies run as one replayed graph (bmi_batch_*): the
    assemblers' plans and matrices, on one context."""
    _free = "bmi_batch_free"

    def __init__(self, assemblers: list):
        super().__init__()
        if not assemblers:
            raise ValueError("empty batch")
        ctx = assemblers[0].ctx
        if any(a.ctx is not ctx for a in assemblers):
            raise ValueError("the assemblers of a batch share one context")
        self.assemblers = list(assemblers)
        n = len(assemblers)
        plans = (ctypes.c_void_p * n)(*[a.plan.h.value for a in assemblers])
        mats = (ctypes.c_void_p * n)(*[a.matrix.h.value for a in assemblers])
        _bmi.check(_bmi.lib().bmi_batch_create(ctx.h, plans, mats, n, ctypes.byref(self.h)))

    def assemble(self) -> None:
        _bmi.check(_bmi.lib().bmi_batch_assemble(self.h))

    def stats(self) -> dict:
        s = _bmi.BatchStats()
        _bmi.check(_bmi.lib().bmi_batch_stats_get(self.h, ctypes.byref(s)))
        return s.as_dict()

    def set_schedule(self, bucket: int, mode: str, wgs_per_cu: int = 0) -> None:
        """Override the tuned far-phase schedule of one bucket: "per-colour",
        "persistent" with 1 .. 3 work-groups per CU, or "per-problem"."""
        if mode not in _bmi.BATCH_MODES:
            raise ValueError(f"mode must be one of {_bmi.BATCH_MODES}")
        _bmi.check(_bmi.lib().bmi_batch_set_schedule(self.h, bucket, _bmi.BATCH_MODES.index(mode), wgs_per_cu))


def assemble_gpu(operator: Operator, test_functions: Space, trial_functions: Space,
                 config: GPUConfiguration | None = None, writeBackStrategy: str | None = None) -> np.ndarray:
    """Dense Galerkin matrix of ``operator`` on (test, trial) as a host
    column-major complex array (ComplexF64 unless config.output == "c32");
    writeBackStrategy "gpu" (default) or "cpu" as in GPUConfiguration."""
    config = dataclasses.replace(config or GPUConfiguration())
    if writeBackStrategy is not None:
        if writeBackStrategy not in ("gpu", "cpu"):
            raise ValueError("writeBackStrategy must be 'gpu' or 'cpu'")
        config.writeBackStrategy = writeBackStrategy
    stream = config.writeBackStrategy == "cpu"
    a = Assembler(operator, test_functions, trial_functions, config, resident=not stream)
    try:
        return a.assemble_to_host() if stream else a.assemble().to_host()
    finally:
        a.close()
