"""Host-side mirror of the reference operator interface over the qclhip C-ABI.

Names, argument meaning and error behaviour follow the production operators
of QCLNEGF.jl (src/numerics/optimized/production_*.jl); fields live in device
memory as torch complex128 tensors whose memory is the reference layout
(NE,Nk,Nb,Nb) column-major, i.e. torch shape (Nb, Nb, Nk, NE) contiguous.
torch supplies device memory and the stream; all arithmetic runs in the HIP
kernels of libqclhip.so.
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import check

_vp = ct.c_void_p
_dp = ct.POINTER(ct.c_double)


def _ptr(t):
    return None if t is None else ct.c_void_p(t.data_ptr())


def _host(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _chost(a):
    """Complex host matrix (column-major) as an interleaved double pointer."""
    a = np.asfortranarray(a, dtype=np.complex128)
    return a, a.ctypes.data_as(_dp)


class Context:
    """One qclhip context bound to a device and to torch's current stream."""

    def __init__(self, device: int = 0):
        lib = _lib.load()
        if not torch.cuda.is_available():
            raise RuntimeError("qclhip needs a ROCm GPU; none is visible")
        self.device = torch.device("cuda", device)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        h = _vp()
        check(lib.qclhip_ctx_create(device, ct.c_void_p(stream), ct.byref(h)), None, "qclhip_ctx_create")
        self.handle = h
        self.lib = lib

    def synchronize(self):
        check(self.lib.qclhip_ctx_synchronize(self.handle), self.handle, "synchronize")

    def close(self):
        if self.handle:
            self.lib.qclhip_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def call(self, name, *args):
        rc = getattr(self.lib, name)(self.handle, *args)
        check(rc, self.handle, name)

    # -- streams, deferred results and graphs --------------------------------
    deferring = False

    def set_stream(self, stream):
        """Enqueue on `stream` (a torch.cuda.Stream) from now on."""
        self.call("qclhip_ctx_set_stream", ct.c_void_p(stream.cuda_stream))

    def defer(self, buffer):
        """Deferred-results mode into the device float64 tensor `buffer`
        (None ends it): scalar-returning calls return Deferred slots."""
        self.call("qclhip_ctx_defer_results", _ptr(buffer), 0 if buffer is None else buffer.numel())
        self.deferring = buffer is not None

    def result_slot(self) -> int:
        return int(self.lib.qclhip_ctx_result_slot(self.handle))

    def graph_begin(self):
        self.call("qclhip_graph_begin")

    def graph_end(self) -> "Graph":
        h = _vp()
        self.call("qclhip_graph_end", ct.byref(h))
        return Graph(self, h)


class Graph:
    """An executable HIP graph captured from a context's stream."""

    def __init__(self, ctx: Context, handle):
        self.ctx, self.handle = ctx, handle

    def launch(self):
        self.ctx.call("qclhip_graph_launch", self.handle)

    def __del__(self):
        try:
            if self.handle:
                self.ctx.lib.qclhip_graph_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


@dataclass
class Deferred:
    """Where a deferred result will land in the context's result buffer."""
    slot: int
    n: int

    def read(self, host_buffer):
        v = host_buffer[self.slot:self.slot + self.n]
        return float(v[0]) if self.n == 1 else np.array(v)


_default = {}


def context(device: int = 0) -> Context:
    if device not in _default:
        _default[device] = Context(device)
    return _default[device]


# ---------------------------------------------------------------- layout helpers

def field_shape(NE, Nk, Nb):
    return (Nb, Nb, Nk, NE)


def empty_field(grid, device=None):
    return torch.empty(field_shape(grid.NE, grid.Nk, grid.Nb), dtype=torch.complex128,
                       device=device or grid.ctx.device)


def zeros_field(grid, device=None):
    t = torch.empty(field_shape(grid.NE, grid.Nk, grid.Nb), dtype=torch.complex128,
                    device=device or grid.ctx.device)
    return fill_zero(t)


def fill_zero(t):
    """t <- 0 by the library's fill kernel (on the context's stream)."""
    ctx = context(t.device.index or 0)
    ctx.call("qclhip_fill_zero", ct.c_int64(t.numel()), _ptr(t))
    return t


def copy_field(src, dst):
    """dst <- src by the library's field kernel."""
    return sum_fields([src], dst)


def to_device(x, device="cuda"):
    """numpy (NE,Nk,Nb,Nb) (any order) -> device tensor in the reference layout."""
    x = np.asarray(x, dtype=np.complex128)
    return torch.from_numpy(np.ascontiguousarray(x.T)).to(device)


def to_host(t):
    """Device tensor in the reference layout -> numpy (NE,Nk,Nb,Nb), Fortran order."""
    return np.asfortranarray(t.detach().cpu().numpy().T)


def h_to_device(h, device="cuda"):
    """numpy h (Nk,Nb,Nb) -> device tensor (Nb,Nb,Nk)."""
    return torch.from_numpy(np.ascontiguousarray(np.asarray(h, np.complex128).T)).to(device)


def h_to_host(t):
    """device h (Nb,Nb,Nk) -> numpy (Nk,Nb,Nb)."""
    return np.ascontiguousarray(t.cpu().numpy().T)


def matrix_to_device(T, device="cuda"):
    """Column-major Nb x Nb complex matrix on the device (torch shape (Nb,Nb) = T^T)."""
    return torch.from_numpy(np.ascontiguousarray(np.asarray(T, np.complex128).T)).to(device)


def _check_field(grid, t, name):
    if t is None:
        return
    if t.dtype != torch.complex128 or not t.is_contiguous():
        raise _lib.ArgumentError(f"{name} must be a contiguous complex128 tensor")
    if tuple(t.shape) != field_shape(grid.NE, grid.Nk, grid.Nb):
        raise _lib.DimensionMismatch(f"{name} has shape {tuple(t.shape)}, expected "
                                     f"{field_shape(grid.NE, grid.Nk, grid.Nb)}")


# ---------------------------------------------------------------- handles

class Grid:
    """ModelGrids on the device (reference/grids.jl:10-47) plus Nb and g_s."""

    def __init__(self, eps, wE, wk, Nb, gs=2, ctx: Context | None = None):
        self.ctx = ctx or context()
        self.eps = np.ascontiguousarray(eps, np.float64)
        self.wE = np.ascontiguousarray(wE, np.float64)
        self.wk = np.ascontiguousarray(wk, np.float64)
        self.NE, self.Nk, self.Nb, self.gs = len(self.eps), len(self.wk), int(Nb), int(gs)
        if len(self.wE) != self.NE:
            raise _lib.DimensionMismatch("energy data differ")
        h = _vp()
        self.ctx.call("qclhip_grid_create", self.NE, self.Nk, self.Nb, self.gs,
                      self.eps.ctypes.data_as(_dp), self.wE.ctypes.data_as(_dp),
                      self.wk.ctypes.data_as(_dp), ct.byref(h))
        self.handle = h
        self.precision, self.engine = "fp64", "vector"

    PRECISION_FP64 = 0
    PRECISION_FP32_REFINE = 1
    PRECISION_FP32_PRODUCTS = 2
    ENGINE_VECTOR = 0
    ENGINE_MATRIX = 1
    ENGINE_FUSED = 2

    def set_precision(self, precision: str = "fp64", refine_steps: int = 2, refine_tol: float = 0.0):
        """GF-solve precision of this grid: "fp64" (the reference's),
        "fp32_refine" (FP32 Dyson factorization + refine_steps FP64 Newton
        steps, FP64 re-solve of blocks whose predicted residual exceeds
        refine_tol; FP64 Keldysh products) or "fp32_products" (the same Dyson
        step with FP32 matrix-core Keldysh products)."""
        code = {"fp64": self.PRECISION_FP64, "fp32_refine": self.PRECISION_FP32_REFINE,
                "fp32_products": self.PRECISION_FP32_PRODUCTS}.get(precision)
        if code is None:
            raise _lib.ArgumentError(f"unknown precision {precision!r}")
        self.ctx.call("qclhip_grid_set_precision", self.handle, code, int(refine_steps), ct.c_double(refine_tol))
        self.precision = precision
        return self

    def set_engine(self, engine: str = "vector"):
        """Engine of the FP64 Keldysh products: "vector" or "matrix" cores, or
        "fused" (the FP64 Dyson step and the vector products in one kernel)."""
        code = {"vector": self.ENGINE_VECTOR, "matrix": self.ENGINE_MATRIX, "fused": self.ENGINE_FUSED}.get(engine)
        if code is None:
            raise _lib.ArgumentError(f"unknown engine {engine!r}")
        self.ctx.call("qclhip_grid_set_engine", self.handle, code)
        self.engine = engine
        return self

    def autotune(self):
        """Pick the faster FP64 Keldysh engine and the tile shapes (energies
        per workgroup) of the FP64 vector Dyson and Keldysh kernels for this
        grid's shape (cached per shape in the context); returns the times
        (ms) and the choices."""
        ms = (ct.c_double * 2)()
        self.ctx.call("qclhip_grid_autotune", self.handle, ms)
        t = (ct.c_double * 8)()
        ch = (ct.c_int * 4)()
        self.ctx.call("qclhip_grid_tuning", self.handle, t, ch)
        self.engine = {0: "vector", 1: "matrix", 2: "fused"}[ch[0]]
        return dict(vector_ms=t[0], matrix_ms=t[1], engine=self.engine, dyson_w4_ms=t[2], dyson_w8_ms=t[3],
                    dyson_waves=ch[1], keldysh_w2_ms=t[4], keldysh_w4_ms=t[5], keldysh_waves=ch[2], fused_ms=t[6],
                    keldysh_split=ch[3], keldysh_split_ms=t[7])

    def set_keldysh_split(self, waves_per_block: int = 0):
        """Waves per block of the FP64 vector Keldysh kernel (0: the rule;
        2/4 at N 25..32, 5 at 33..40, 2/3/6 at 41..48); bitwise the same
        results."""
        self.ctx.call("qclhip_grid_set_keldysh_split", self.handle, waves_per_block)
        return self

    def set_tiles(self, dyson_waves: int = 0, keldysh_waves: int = 0):
        """Energies per workgroup of the FP64 vector Dyson (4, 8) and Keldysh
        (2, 4) kernels; 0 restores the built-in rule."""
        self.ctx.call("qclhip_grid_set_tiles", self.handle, dyson_waves, keldysh_waves)
        return self

    def refine_fallbacks(self) -> int:
        """Blocks of the last mixed-precision Dyson call re-solved in FP64."""
        n = ct.c_int64()
        self.ctx.call("qclhip_grid_refine_fallbacks", self.handle, ct.byref(n))
        return n.value

    @property
    def nfield(self):
        return self.NE * self.Nk * self.Nb * self.Nb

    @property
    def ncols(self):
        return self.Nk * self.Nb * self.Nb

    def __del__(self):
        try:
            if self.handle:
                self.ctx.lib.qclhip_grid_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class EnergyShiftPlan:
    """build_shift_plan(eps, delta) on the device (production_cache.jl:258-288)."""

    def __init__(self, grid: Grid, delta: float):
        self.grid, self.delta = grid, float(delta)
        h = _vp()
        grid.ctx.call("qclhip_shift_create", grid.handle, ct.c_double(delta), ct.byref(h))
        self.handle = h

    def __del__(self):
        try:
            if self.handle:
                self.grid.ctx.lib.qclhip_shift_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


_KINDS = {"dense": _lib.KERNEL_DENSE, "momentum_independent": _lib.KERNEL_MOMENTUM_INDEPENDENT,
          "separable_ifr": _lib.KERNEL_SEPARABLE_IFR, "low_rank": _lib.KERNEL_LOW_RANK}


class ProductionKernel:
    """A scattering-kernel representation (production_contracts.jl:262-323).

    dense: a = Kflat (Nc x Nc); momentum_independent: a = block (Nb^2 x Nb^2);
    separable_ifr: a = left, b = right, c = radial; low_rank: a = left,
    b = right_adjoint. Host matrices, column-major semantics (numpy arrays in
    natural (row, column) indexing).
    """

    def __init__(self, grid: Grid, kind: str, a=None, b=None, c=None, rank: int = 0, _handle=None):
        self.grid, self.kind, self.rank = grid, kind, int(rank)
        if _handle is not None:
            self.handle = _handle
            return
        keep = [_chost(x) if x is not None else (None, None) for x in (a, b, c)]
        h = _vp()
        grid.ctx.call("qclhip_kernel_create", grid.handle, _KINDS[kind], self.rank,
                      keep[0][1], keep[1][1], keep[2][1], ct.byref(h))
        self.handle = h

    @classmethod
    def dense_from_device(cls, grid: Grid, Kt) -> "ProductionKernel":
        """The dense representation from device memory: Kt is a contiguous
        complex128 tensor of shape (Nc, Nc) with Kt[q', q] = Kflat[q, q'] (the
        memory of the column-major Kflat); qclhip_kernel_create_dense_device
        folds the weights into the library's own copy."""
        Nc = grid.Nk * grid.Nb * grid.Nb
        if Kt.dtype != torch.complex128 or tuple(Kt.shape) != (Nc, Nc) or not Kt.is_contiguous():
            raise _lib.DimensionMismatch("Kflat must be a contiguous complex128 (Nc, Nc) device tensor")
        h = _vp()
        grid.ctx.call("qclhip_kernel_create_dense_device", grid.handle, _ptr(Kt), ct.byref(h))
        return cls(grid, "dense", _handle=h)

    def adjoint(self) -> "ProductionKernel":
        """(K diag(wk))^dagger in the same representation (qclhip_kernel_adjoint)."""
        h = _vp()
        self.grid.ctx.call("qclhip_kernel_adjoint", self.handle, ct.byref(h))
        return ProductionKernel(self.grid, self.kind, rank=self.rank, _handle=h)

    def __del__(self):
        try:
            if self.handle:
                self.grid.ctx.lib.qclhip_kernel_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class FFTHilbertPlan:
    """ProductionFFTHilbertPlan (fft_parallel.jl:27-143) on the device."""

    def __init__(self, grid: Grid, ncols: int | None = None, verify_roundoff: bool = True):
        self.grid = grid
        self.ncols = int(ncols if ncols is not None else grid.ncols)
        self.verify = bool(verify_roundoff)
        h = _vp()
        grid.ctx.call("qclhip_hilbert_create", grid.handle, ct.c_int64(self.ncols), int(self.verify),
                      ct.byref(h))
        self.handle = h

    PATHS = {"auto": 0, "lds": 1, "hbm": 2}

    def set_multiplicity(self, mult: int):
        """Each column stands for mult identical columns in the round-off
        witness (a momentum-independent self-energy transformed once for
        all Nk momenta)."""
        self.grid.ctx.call("qclhip_hilbert_set_multiplicity", self.handle, ct.c_int64(int(mult)))
        return self

    def set_path(self, path: str):
        """'auto' (by L), 'lds' (one workgroup per column, L <= 16384) or
        'hbm' (the four-step transform through HBM scratch, L >= 4096)."""
        self.grid.ctx.call("qclhip_hilbert_set_path", self.handle, self.PATHS[path])
        return self

    @property
    def path(self) -> str:
        v = ct.c_int(0)
        self.grid.ctx.call("qclhip_hilbert_get_path", self.handle, ct.byref(v))
        return {1: "lds", 2: "hbm"}[v.value]

    def __del__(self):
        try:
            if self.handle:
                self.grid.ctx.lib.qclhip_hilbert_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


# ---------------------------------------------------------------- operators

def retarded_green_production(grid: Grid, h, sigR=None, eta: float = 0.0, GR=None, scale=None,
                              check_status: bool = True):
    """_retarded_green_production (production_operators.jl:1337-1385).

    Returns (GR, scale). Raises SingularException for a singular block."""
    _check_field(grid, sigR, "sigmaR")
    GR = empty_field(grid) if GR is None else GR
    scale = torch.empty((grid.Nk, grid.NE), dtype=torch.float64, device=grid.ctx.device) \
        if scale is None else scale
    grid.ctx.call("qclhip_retarded_green", grid.handle, _ptr(h), _ptr(sigR), ct.c_double(eta),
                  _ptr(GR), _ptr(scale))
    if check_status:
        grid.ctx.synchronize()
    return GR, scale


def dyson_keldysh(grid: Grid, h, sigR, sigL, sigG, GR=None, GL=None, GG=None, scale=None,
                  traces=None, check_status: bool = False):
    """Fused Dyson plus raw Keldysh products G^R S G^R^dagger (production_state.jl:14-48)."""
    for t, n in ((sigR, "sigmaR"), (sigL, "sigmaL"), (sigG, "sigmaG")):
        _check_field(grid, t, n)
    GR = empty_field(grid) if GR is None else GR
    GL = empty_field(grid) if GL is None else GL
    GG = empty_field(grid) if GG is None else GG
    grid.ctx.call("qclhip_dyson_keldysh", grid.handle, _ptr(h), _ptr(sigR), _ptr(sigL), _ptr(sigG),
                  ct.c_double(0.0), _ptr(GR), _ptr(GL), _ptr(GG), _ptr(scale), _ptr(traces))
    if check_status:
        grid.ctx.synchronize()
    return GR, GL, GG


def strided_field(grid: Grid, plane: int):
    """A zeroed field whose (a,b) planes lie `plane` elements apart: the
    (Nb, Nb, Nk, NE) view of (Nb*Nb, plane) device storage."""
    if plane < grid.NE * grid.Nk:
        raise _lib.ArgumentError("the plane stride must be at least NE*Nk")
    store = torch.zeros((grid.Nb * grid.Nb, plane), dtype=torch.complex128, device=grid.ctx.device)
    return store[:, :grid.NE * grid.Nk].view(field_shape(grid.NE, grid.Nk, grid.Nb))


def dyson_keldysh_strided(grid: Grid, plane: int, h, sigR, sigL, sigG, GR, GL, GG, scale=None, traces=None):
    """The Dyson + Keldysh launch pair of dyson_keldysh on plane-strided fields
    (views from strided_field with the same plane)."""
    want = (grid.Nb * plane, plane, grid.NE, 1)
    for t, n in ((sigR, "sigmaR"), (sigL, "sigmaL"), (sigG, "sigmaG"), (GR, "GR"), (GL, "GL"), (GG, "GG")):
        if (t.dtype != torch.complex128 or t.device != grid.ctx.device
                or tuple(t.shape) != field_shape(grid.NE, grid.Nk, grid.Nb) or t.stride() != want):
            raise _lib.DimensionMismatch(f"{n}: expected a complex128 field with strides {want}")
    grid.ctx.call("qclhip_dyson_keldysh_strided", grid.handle, ct.c_int64(plane), _ptr(h), _ptr(sigR),
                  _ptr(sigL), _ptr(sigG), ct.c_double(0.0), _ptr(GR), _ptr(GL), _ptr(GG), _ptr(scale),
                  _ptr(traces))
    return GR, GL, GG


@dataclass
class GreenState:
    """GreenState (types.jl:733-740) with device fields."""
    GR: torch.Tensor
    GL: torch.Tensor
    GG: torch.Tensor
    A: torch.Tensor | None
    scale: torch.Tensor


def current_green_production(grid: Grid, h, families, target: float, state: GreenState | None = None,
                             spectral: bool = True, return_scalars: bool = True, totals=None,
                             occupation_normalization: str = "paired_convex", allocation_free: bool = False,
                             raw_screen=None):
    """_current_green_production (production_state.jl:1-74).

    families: list of (SR, SL, SG) device triples, the embedding total first,
    then the scattering mechanisms in mechanism order. Returns
    (GreenState, info) where info holds lambda, number, holes and the paired
    coefficients (a, c) when return_scalars (the call then synchronizes).
    totals: optional (SR, SL, SG) fields that receive the summed families.
    occupation_normalization: "paired_convex" (default) or "scalar_lesser",
    the reference's legacy branch G< = G<raw/lambda, G> = G^R - G^A + G<
    (production_state.jl:54-61; the options of algorithm_options.jl:63).
    allocation_free: use qclhip_current_green_noalloc (a fixed device scratch,
    no device allocation; NE*Nk <= 524,288, at most 8 families).
    raw_screen: optional uint8 device tensor of NE*Nk entries that receives
    the raw-product screen words of qclhip_current_green_screened (FP64
    vector engine), for residual_suite/residual_state(raw_screen=...)."""
    if occupation_normalization not in ("paired_convex", "scalar_lesser"):
        raise _lib.ArgumentError("occupation_normalization must be :paired_convex or :scalar_lesser")
    scalar = occupation_normalization == "scalar_lesser"
    if scalar and totals is not None:
        raise _lib.ArgumentError("totals are written by the paired_convex path only")
    if allocation_free and (scalar or totals is not None):
        raise _lib.ArgumentError("the allocation-free call is the plain paired_convex current_green")
    if raw_screen is not None:
        if scalar or allocation_free:
            raise _lib.ArgumentError("raw-product screens come with the paired_convex current_green")
        _check_screen(grid, raw_screen)
    nf = len(families)
    for fam in families:
        for t in fam:
            _check_field(grid, t, "self-energy")
    if state is None:
        state = GreenState(empty_field(grid), empty_field(grid), empty_field(grid),
                           empty_field(grid) if spectral else None,
                           torch.empty((grid.Nk, grid.NE), dtype=torch.float64, device=grid.ctx.device))
    arr = lambda k: (_vp * nf)(*[f[k].data_ptr() for f in families])
    out = (ct.c_double * 5)()
    slot0 = grid.ctx.result_slot() if grid.ctx.deferring else 0
    common = (grid.handle, _ptr(h), nf, arr(0), arr(1), arr(2), ct.c_double(target), _ptr(state.GR),
              _ptr(state.GL), _ptr(state.GG), _ptr(state.A if spectral else None), _ptr(state.scale),
              out if return_scalars else None)
    if raw_screen is not None:
        if totals is not None:
            for t in totals:
                _check_field(grid, t, "total")
        tp = [_ptr(t) for t in totals] if totals is not None else [None, None, None]
        grid.ctx.call("qclhip_current_green_screened", *common, *tp, _ptr(raw_screen))
    elif allocation_free:
        grid.ctx.call("qclhip_current_green_noalloc", *common)
    elif scalar:
        grid.ctx.call("qclhip_current_green_scalar_lesser", *common)
    elif totals is None:
        grid.ctx.call("qclhip_current_green", *common)
    else:
        for t in totals:
            _check_field(grid, t, "total")
        grid.ctx.call("qclhip_current_green_totals", *common, *(_ptr(t) for t in totals))
    if return_scalars and grid.ctx.deferring:
        return state, Deferred(slot0, 5)
    info = dict(lam=out[0], number=out[1], holes=out[2], a=out[3], c=out[4]) if return_scalars else None
    return state, info


def _check_screen(grid: Grid, raw_screen):
    if (not isinstance(raw_screen, torch.Tensor) or raw_screen.dtype != torch.uint8 or not raw_screen.is_cuda
            or not raw_screen.is_contiguous() or raw_screen.numel() != grid.NE * grid.Nk):
        raise _lib.DimensionMismatch("raw_screen must be a contiguous uint8 device tensor of NE*Nk entries")


def screen_buffer(grid: Grid):
    """Device storage of one raw-product screen word per (e,m) block."""
    return torch.empty(grid.NE * grid.Nk, dtype=torch.uint8, device=grid.ctx.device)


def number_functional(grid: Grid, G) -> float:
    """_number_functional_production (production_operators.jl:1417-1445)."""
    v = ct.c_double()
    slot0 = grid.ctx.result_slot() if grid.ctx.deferring else 0
    grid.ctx.call("qclhip_number_functional", grid.handle, _ptr(G), ct.byref(v))
    return Deferred(slot0, 1) if grid.ctx.deferring else v.value


def spectral_function(grid: Grid, GR, A=None):
    A = empty_field(grid) if A is None else A
    grid.ctx.call("qclhip_spectral", grid.handle, _ptr(GR), _ptr(A))
    return A


def normalize_keldysh_pair(GL, GG, a: float, c: float):
    ctx = context(GL.device.index or 0)
    ctx.call("qclhip_normalize_pair", ct.c_int64(GL.numel()), _ptr(GL), _ptr(GG), ct.c_double(a),
             ct.c_double(c))


def apply_energy_shift(plan: EnergyShiftPlan, X, Y=None):
    """apply_energy_shift! (production_cache.jl:300-325)."""
    Y = torch.empty_like(X) if Y is None else Y
    ncols = X.numel() // plan.grid.NE
    plan.grid.ctx.call("qclhip_apply_shift", plan.handle, ct.c_int64(ncols), _ptr(X), _ptr(Y))
    return Y


def production_static_contraction(kernel: ProductionKernel, G, qK: float = 1.0, alpha: float = 1.0,
                                  beta: float = 0.0, S=None):
    """_production_static_contraction! (production_operators.jl:402-658)."""
    _check_field(kernel.grid, G, "G")
    S = empty_field(kernel.grid) if S is None else S
    kernel.grid.ctx.call("qclhip_contract", kernel.handle, _ptr(G), ct.c_double(qK), ct.c_double(alpha),
                         ct.c_double(beta), _ptr(S))
    return S


def compact_field(grid, device=None):
    """One momentum's columns (NE x Nb^2) of a momentum-independent field."""
    return torch.empty(field_shape(grid.NE, 1, grid.Nb), dtype=torch.complex128,
                       device=device or grid.ctx.device)


def _check_compact(grid, t, name):
    if t.dtype != torch.complex128 or not t.is_contiguous():
        raise _lib.ArgumentError(f"{name} must be a contiguous complex128 tensor")
    if tuple(t.shape) != field_shape(grid.NE, 1, grid.Nb):
        raise _lib.DimensionMismatch(f"{name} has shape {tuple(t.shape)}, expected "
                                     f"{field_shape(grid.NE, 1, grid.Nb)}")


def momentum_independent_contraction_compact(kernel: ProductionKernel, G, qK: float = 1.0, alpha: float = 1.0,
                                             Y=None):
    """The momentum-independent branch of _production_static_contraction!
    (production_operators.jl:535-595) without its broadcast: the NE x Nb^2
    columns that every momentum of the result holds (bit for bit)."""
    _check_field(kernel.grid, G, "G")
    Y = compact_field(kernel.grid) if Y is None else Y
    _check_compact(kernel.grid, Y, "Y")
    kernel.grid.ctx.call("qclhip_contract_compact", kernel.handle, _ptr(G), ct.c_double(qK), ct.c_double(alpha),
                         _ptr(Y))
    return Y


def broadcast_momenta(grid: Grid, Y, S=None):
    """S(e, m, a, b) = Y(e, a, b) for every momentum m (the broadcast of the
    momentum-independent contraction)."""
    _check_compact(grid, Y, "Y")
    S = empty_field(grid) if S is None else S
    _check_field(grid, S, "S")
    grid.ctx.call("qclhip_broadcast_momenta", grid.handle, _ptr(Y), _ptr(S))
    return S


def production_lo_contraction(kernel: ProductionKernel, GL, GG, qK: float, NLO: float,
                              plus: EnergyShiftPlan, minus: EnergyShiftPlan, SL=None, SG=None):
    """_production_lo_contraction with thermal N_LO (production_operators.jl:1089-1162)."""
    SL = empty_field(kernel.grid) if SL is None else SL
    SG = empty_field(kernel.grid) if SG is None else SG
    kernel.grid.ctx.call("qclhip_lo_contraction", kernel.handle, _ptr(GL), _ptr(GG), ct.c_double(qK),
                         ct.c_double(NLO), plus.handle, minus.handle, _ptr(SL), _ptr(SG))
    return SL, SG


def production_fft_hilbert_transform(plan: FFTHilbertPlan, Gamma, return_roundoff: bool = False,
                                      Lambda=None):
    """production_fft_hilbert_transform (fft_parallel.jl:247-291)."""
    Lambda = torch.empty_like(Gamma) if Lambda is None else Lambda
    r = ct.c_double(0.0)
    plan.grid.ctx.call("qclhip_fft_hilbert", plan.handle, _ptr(Gamma), _ptr(Lambda),
                       ct.byref(r) if return_roundoff else None)
    return (Lambda, r.value) if return_roundoff else Lambda


def retarded_self_energy_fft(plan: FFTHilbertPlan, SL, SG, return_roundoff: bool = False, SR=None,
                             Gamma=None):
    """_retarded_self_energy_fft_production (production_operators.jl:842-884).

    Returns (SR, Gamma[, roundoff]); Gamma is only materialized when passed."""
    SR = torch.empty_like(SL) if SR is None else SR
    r = ct.c_double(0.0)
    ctx = plan.grid.ctx
    slot0 = ctx.result_slot() if ctx.deferring else 0
    ctx.call("qclhip_retarded_self_energy", plan.handle, _ptr(SL), _ptr(SG), _ptr(SR),
             _ptr(Gamma), ct.byref(r) if return_roundoff else None)
    if return_roundoff:
        return SR, Gamma, (Deferred(slot0, 1) if ctx.deferring else r.value)
    return SR, Gamma


def embedding_component_production(grid: Grid, T, X, plan: EnergyShiftPlan, S=None):
    """_embedding_component_production (production_operators.jl:886-922); T on the device."""
    S = empty_field(grid) if S is None else S
    grid.ctx.call("qclhip_embedding_component", grid.handle, _ptr(T), _ptr(X), plan.handle, _ptr(S))
    return S


def sum_fields(xs, out):
    ctx = context(out.device.index or 0)
    arr = (_vp * len(xs))(*[x.data_ptr() for x in xs])
    ctx.call("qclhip_sum_fields", ct.c_int64(out.numel()), len(xs), arr, _ptr(out))
    return out


def mix_family_production(old, cand, alpha: float):
    """_mix_family_production! for one component (production_operators.jl:1475-1491)."""
    ctx = context(old.device.index or 0)
    ctx.call("qclhip_mix", ct.c_int64(old.numel()), _ptr(old), _ptr(cand), ct.c_double(alpha))
    return old


def mix_fields(pairs, alpha: float):
    """Mix every (old, candidate) pair in one pass: old <- (1-alpha) old + alpha cand."""
    if not pairs:
        return
    ctx = context(pairs[0][0].device.index or 0)
    n = pairs[0][0].numel()
    olds = (_vp * len(pairs))(*[o.data_ptr() for o, _ in pairs])
    cands = (_vp * len(pairs))(*[c.data_ptr() for _, c in pairs])
    ctx.call("qclhip_mix_fields", ct.c_int64(n), len(pairs), olds, cands, ct.c_double(alpha))


def mix_fields_residual(pairs, alpha: float, partner=None, totals=None, compact=None, grid=None):
    """Mix every (old, candidate) pair in one pass and return the self-energy
    residual ratios of the unmixed pairs: one per pair, then one per couple
    (p, partner[p]) with partner[p] > p, for the sum of the two pairs. A
    Deferred of that many values in deferred-results mode.

    totals: optional list of (field, [pair indices]) or (field, [pair
    indices], flags); each field receives the left-fold sum of its members'
    mixed values (every pair in one total). flags: "accumulate" continues
    the fold from the field's current value; "candidate" sums the members'
    candidate values instead (the pairs then also belong to a mixing total);
    both may be given as "accumulate,candidate".

    compact: optional per-pair flags (with the grid and totals): a flagged
    pair's candidate is one momentum's columns (compact_field), read for
    every momentum (bitwise the mixing of its broadcast field)."""
    pairs = list(pairs)
    if not pairs:
        raise _lib.ArgumentError("no pairs to mix")
    n = pairs[0][0].numel()
    cflags = [0] * len(pairs) if compact is None else [int(bool(x)) for x in compact]
    if compact is not None and (grid is None or totals is None or len(cflags) != len(pairs)):
        raise _lib.ArgumentError("compact candidates need the grid, the totals and one flag per pair")
    for (o, c), cf in zip(pairs, cflags):
        nc = n // grid.Nk if cf else n
        if o.dtype != torch.complex128 or c.dtype != torch.complex128 or o.numel() != n or c.numel() != nc \
                or not o.is_contiguous() or not c.is_contiguous():
            raise _lib.ArgumentError("mixing pairs must be contiguous complex128 fields of one size")
    partner = [-1] * len(pairs) if partner is None else list(partner)
    nterms = len(pairs) + sum(1 for p, q in enumerate(partner) if q > p)
    ctx = context(pairs[0][0].device.index or 0)
    olds = (_vp * len(pairs))(*[o.data_ptr() for o, _ in pairs])
    cands = (_vp * len(pairs))(*[c.data_ptr() for _, c in pairs])
    out = (ct.c_double * nterms)()
    slot0 = ctx.result_slot() if ctx.deferring else 0
    if totals is None:
        ctx.call("qclhip_mix_fields_residual", ct.c_int64(n), len(pairs), olds, cands, ct.c_double(alpha),
                 (ct.c_int * len(pairs))(*partner), out)
    else:
        totals = [tuple(t) if len(t) == 3 else (t[0], t[1], "") for t in totals]
        for t, _, _ in totals:
            if t.dtype != torch.complex128 or t.numel() != n or not t.is_contiguous():
                raise _lib.ArgumentError("totals must be contiguous complex128 fields of the pairs' size")
        counts = [len(m) for _, m, _ in totals]
        members = [int(p) for _, m, _ in totals for p in m]
        flag_bits = {"": 0, "accumulate": 1, "candidate": 2}
        try:
            flags = [sum(flag_bits[f.strip()] for f in fl.split(",")) for _, _, fl in totals]
        except KeyError:
            raise _lib.ArgumentError("total flags are 'accumulate' and/or 'candidate'") from None
        common = (ct.c_int64(n), len(pairs), olds, cands, ct.c_double(alpha), (ct.c_int * len(pairs))(*partner),
                  len(totals), (ct.c_int * len(counts))(*counts), (ct.c_int * len(members))(*members))
        tptr = (_vp * len(totals))(*[t.data_ptr() for t, _, _ in totals])
        if compact is not None:
            ctx.call("qclhip_mix_fields_residual_compact", grid.handle, *common, (ct.c_int * len(flags))(*flags),
                     tptr, (ct.c_int * len(cflags))(*cflags), out)
        elif any(flags):
            ctx.call("qclhip_mix_fields_residual_accumulate", *common, (ct.c_int * len(flags))(*flags), tptr, out)
        else:
            ctx.call("qclhip_mix_fields_residual_totals", *common, tptr, out)
    return Deferred(slot0, nterms) if ctx.deferring else list(out)


def mix_fields_change(pairs, alpha: float):
    """Mix every (old, candidate) pair in one pass; returns
    (sum ||c - o||^2, sum ||c||^2) over the unmixed pairs."""
    pairs = list(pairs)
    n = pairs[0][0].numel()
    ctx = context(pairs[0][0].device.index or 0)
    olds = (_vp * len(pairs))(*[o.data_ptr() for o, _ in pairs])
    cands = (_vp * len(pairs))(*[c.data_ptr() for _, c in pairs])
    out = (ct.c_double * 2)()
    ctx.call("qclhip_mix_fields_change", ct.c_int64(n), len(pairs), olds, cands, ct.c_double(alpha), out)
    return out[0], out[1]


def boundary_flux_bar(grid: Grid, plusL, plusG, GL, GG) -> float:
    """_production_boundary_flux_bar (production_iteration.jl:189-201)."""
    v = ct.c_double()
    slot0 = grid.ctx.result_slot() if grid.ctx.deferring else 0
    grid.ctx.call("qclhip_boundary_flux", grid.handle, _ptr(plusL), _ptr(plusG), _ptr(GL), _ptr(GG),
                  ct.byref(v))
    return Deferred(slot0, 1) if grid.ctx.deferring else v.value


def state_populations(grid: Grid, GL, target: float):
    """_production_state_populations (production_iteration.jl:204-215)."""
    out = np.zeros(grid.Nb)
    slot0 = grid.ctx.result_slot() if grid.ctx.deferring else 0
    grid.ctx.call("qclhip_state_populations", grid.handle, _ptr(GL), ct.c_double(target),
                  out.ctypes.data_as(_dp))
    return Deferred(slot0, grid.Nb) if grid.ctx.deferring else out


def seed_number_weights(grid: Grid, GR):
    """_seed_number_weights_production (seed.jl:28-66); host vector q[NE]."""
    q = np.zeros(grid.NE)
    grid.ctx.call("qclhip_seed_weights", grid.handle, _ptr(GR), q.ctypes.data_as(_dp))
    return q


def seed_lesser(grid: Grid, GR, mu: float, kBT: float, GL=None):
    """_seed_lesser_production! (seed.jl:99-128)."""
    GL = empty_field(grid) if GL is None else GL
    grid.ctx.call("qclhip_seed_lesser", grid.handle, _ptr(GR), ct.c_double(mu), ct.c_double(kBT), _ptr(GL))
    return GL


def fermi(e, mu, kBT):
    """_fermi (reference/scba.jl:1-6), vectorized."""
    x = (np.asarray(e, np.float64) - mu) / kBT
    with np.errstate(over="ignore"):
        f = 1.0 / (np.exp(x) + 1.0)
    return np.where(x > 40.0, 0.0, np.where(x < -40.0, 1.0, f))


def seed_chemical_potential(eps, q, kBT: float, target: float) -> float:
    """Bisection of sum_e q_e f(e; mu) = target on [e1 - 50kT, eN + 50kT]
    (seed.jl:175-182), down to adjacent floating-point numbers."""
    number = lambda mu: float(np.sum(q * fermi(eps, mu, kBT))) - target
    lo, hi = eps[0] - 50 * kBT, eps[-1] + 50 * kBT
    flo, fhi = number(lo), number(hi)
    if not flo * fhi <= 0:
        raise _lib.ArgumentError("seed chemical-potential root is not bracketed")
    if flo == 0:
        return lo
    if fhi == 0:
        return hi
    for _ in range(4000):
        mid = lo + 0.5 * (hi - lo)
        if mid <= lo or mid >= hi:
            break
        fm = number(mid)
        if fm == 0:
            return mid
        if (fm < 0) == (flo < 0):
            lo, flo = mid, fm
        else:
            hi, fhi = mid, fm
    return lo if abs(flo) <= abs(fhi) else hi


def seed_green_production(grid: Grid, h, eta: float, kBT: float, target: float, spectral: bool = True):
    """_seed_green_production (seed.jl:151-197): Dyson with eta and Sigma = 0,
    spectral number weights, bisection for mu, G< = -f (G^R - G^R^dagger)."""
    GR, scale = retarded_green_production(grid, h, None, eta)
    q = seed_number_weights(grid, GR)
    mu = seed_chemical_potential(grid.eps, q, kBT, target)
    GL = seed_lesser(grid, GR, mu, kBT)
    GG, A = green_state(grid, GR, GL, spectral=spectral)
    return GreenState(GR, GL, GG, A, scale), mu


def green_state(grid: Grid, GR, GL, spectral: bool = True, GG=None, A=None):
    """G> = G^R - G^A + G< and A = i(G^R - G^A) (production_operators.jl:1387-1415)."""
    GG = empty_field(grid) if GG is None else GG
    A = (empty_field(grid) if A is None else A) if spectral else None
    grid.ctx.call("qclhip_green_state", grid.handle, _ptr(GR), _ptr(GL), _ptr(GG), _ptr(A))
    return GG, A


# ---- A19: adjoint (vector-Jacobian products); no reference counterpart ----
# Gradient convention: Zbar = dJ/dRe Z + i dJ/dIm Z for a real functional J.

def current_green_adjoint(grid: Grid, families, target: float, info: dict, green: GreenState, GLbar, GGbar,
                          GRbar=None, want_hbar: bool = True, out=None):
    """Reverse of current_green_production: gradients (SRbar, SLbar, SGbar) of
    the total self-energy (= of every family) and hbar (Nb,Nb,Nk) from the
    gradients of the normalized outputs. `info` is the forward's scalars."""
    nf = len(families)
    arr = lambda k: (_vp * nf)(*[f[k].data_ptr() for f in families])
    SRb, SLb, SGb = out if out is not None else (empty_field(grid), empty_field(grid), empty_field(grid))
    hbar = torch.empty((grid.Nb, grid.Nb, grid.Nk), dtype=torch.complex128, device=grid.ctx.device) \
        if want_hbar else None
    coef = (ct.c_double * 5)(info["lam"], info["number"], info["holes"], info["a"], info["c"])
    grid.ctx.call("qclhip_current_green_adjoint", grid.handle, nf, arr(1), arr(2), ct.c_double(target), coef,
                  _ptr(green.GR), _ptr(green.GL), _ptr(green.GG), _ptr(GRbar), _ptr(GLbar), _ptr(GGbar),
                  _ptr(SRb), _ptr(SLb), _ptr(SGb), _ptr(hbar))
    return (SRb, SLb, SGb), hbar


def boundary_flux_adjoint(grid: Grid, Jbar: float, plusL, plusG, GL, GG, plusLbar=None, plusGbar=None,
                          GLbar=None, GGbar=None):
    """Reverse of boundary_flux_bar; accumulates into the given gradients."""
    grid.ctx.call("qclhip_boundary_flux_adjoint", grid.handle, ct.c_double(Jbar), _ptr(plusL), _ptr(plusG),
                  _ptr(GL), _ptr(GG), _ptr(plusLbar), _ptr(plusGbar), _ptr(GLbar), _ptr(GGbar))


def shift_transpose(plan: EnergyShiftPlan, Y, X=None, alpha: float = 1.0, accumulate: bool = False):
    """X (+)= alpha S^T Y, the transpose of apply_energy_shift."""
    X = torch.empty_like(Y) if X is None else X
    plan.grid.ctx.call("qclhip_shift_transpose", plan.grid.handle, plan.handle, _ptr(Y), ct.c_double(alpha),
                       int(accumulate), _ptr(X))
    return X


def retarded_self_energy_adjoint(plan: FFTHilbertPlan, SRbar, SLbar, SGbar):
    """Reverse of retarded_self_energy_fft; accumulates into SLbar, SGbar."""
    plan.grid.ctx.call("qclhip_retarded_self_energy_adjoint", plan.handle, _ptr(SRbar), _ptr(SLbar), _ptr(SGbar))


def embedding_adjoint(grid: Grid, T, plan: EnergyShiftPlan, Sbar, Xbar):
    """Reverse of embedding_component_production; accumulates
    Xbar += S^T (T^dagger Sbar T)."""
    grid.ctx.call("qclhip_embedding_adjoint", grid.handle, _ptr(T), plan.handle, _ptr(Sbar), _ptr(Xbar))


def antihermitian_part(grid: Grid, X):
    """X_em <- (X_em - X_em^dagger)/2 in place."""
    grid.ctx.call("qclhip_antihermitian_part", grid.handle, _ptr(X))
    return X


def residual_suite(grid: Grid, h, green, families, candidate_lesser, raw_screen=None):
    """production_residual_suite (residuals.jl:537-772) on the device.

    green: a GreenState (A may be None: formed from G^R in the kernel,
    bitwise the spectral function); families: (SR, SL, SG) triples of the
    total self-energy; candidate_lesser: the unmixed candidate S< as a list of
    fields summed in order. raw_screen: the screen words of the
    current_green_production(raw_screen=...) call that produced green from
    these families' total (qclhip_residual_suite_prescreened: the same
    results). Returns dict r_D, r_A, r_K, r_PSD, r_caus, or a Deferred of 5
    values in deferred-results mode."""
    if raw_screen is not None:
        _check_screen(grid, raw_screen)
    for fam in families:
        for t in fam:
            _check_field(grid, t, "self-energy")
    cands = list(candidate_lesser)
    for t in cands:
        _check_field(grid, t, "candidate")
    nf = len(families)
    arr = lambda k: (_vp * nf)(*[f[k].data_ptr() for f in families])
    carr = (_vp * len(cands))(*[t.data_ptr() for t in cands])
    out = (ct.c_double * 5)()
    slot0 = grid.ctx.result_slot() if grid.ctx.deferring else 0
    if raw_screen is None:
        grid.ctx.call("qclhip_residual_suite", grid.handle, _ptr(h), _ptr(green.GR), _ptr(green.GL),
                      _ptr(green.GG), _ptr(green.A), nf, arr(0), arr(1), arr(2), len(cands), carr, out)
    else:
        grid.ctx.call("qclhip_residual_suite_prescreened", grid.handle, _ptr(h), _ptr(green.GR), _ptr(green.GL),
                      _ptr(green.GG), _ptr(green.A), nf, arr(0), arr(1), arr(2), len(cands), carr,
                      _ptr(raw_screen), out)
    if grid.ctx.deferring:
        return Deferred(slot0, 5)
    return dict(r_D=out[0], r_A=out[1], r_K=out[2], r_PSD=out[3], r_caus=out[4])


def residual_state(grid: Grid, h, green, families, raw_screen=None):
    """Every term of production_residual_suite (residuals.jl:537-772) but r_K:
    r_D, r_A, r_PSD, r_caus from the Green state (A may be None: formed from
    G^R) and the total self-energy, before the candidates exist; raw_screen
    as for residual_suite. Returns the dict of the five keys with r_K = 0, or
    a Deferred of 5 values."""
    if raw_screen is not None:
        _check_screen(grid, raw_screen)
    for fam in families:
        for t in fam:
            _check_field(grid, t, "self-energy")
    nf = len(families)
    arr = lambda k: (_vp * nf)(*[f[k].data_ptr() for f in families])
    out = (ct.c_double * 5)()
    slot0 = grid.ctx.result_slot() if grid.ctx.deferring else 0
    if raw_screen is None:
        grid.ctx.call("qclhip_residual_state", grid.handle, _ptr(h), _ptr(green.GR), _ptr(green.GL),
                      _ptr(green.GG), _ptr(green.A), nf, arr(0), arr(1), arr(2), out)
    else:
        grid.ctx.call("qclhip_residual_state_prescreened", grid.handle, _ptr(h), _ptr(green.GR), _ptr(green.GL),
                      _ptr(green.GG), _ptr(green.A), nf, arr(0), arr(1), arr(2), _ptr(raw_screen), out)
    if grid.ctx.deferring:
        return Deferred(slot0, 5)
    return dict(r_D=out[0], r_A=out[1], r_K=out[2], r_PSD=out[3], r_caus=out[4])


def residual_keldysh(grid: Grid, green, candidate_lesser) -> float:
    """r_K of production_residual_suite (residuals.jl:643-664) alone, from
    G^R, G< and the unmixed candidate S< (a list of fields summed in order).
    Returns r_K, or a Deferred of 5 values (r_K in position 2)."""
    cands = list(candidate_lesser)
    for t in cands:
        _check_field(grid, t, "candidate")
    carr = (_vp * len(cands))(*[t.data_ptr() for t in cands])
    out = (ct.c_double * 5)()
    slot0 = grid.ctx.result_slot() if grid.ctx.deferring else 0
    grid.ctx.call("qclhip_residual_keldysh", grid.handle, _ptr(green.GR), _ptr(green.GL), len(cands), carr, out)
    if grid.ctx.deferring:
        return Deferred(slot0, 5)
    return out[2]


def selfenergy_residual(old, candidate):
    """production_selfenergy_residual (residuals.jl:802-865): old and
    candidate are (SR, SL, SG) triples of equal shapes."""
    old, candidate = list(old), list(candidate)
    if len(old) != 3 or len(candidate) != 3:
        raise _lib.ArgumentError("self-energy residual takes (R, <, >) triples")
    n = old[0].numel()
    for t in old + candidate:
        if t.dtype != torch.complex128 or t.numel() != n or not t.is_contiguous() or t.device != old[0].device:
            raise _lib.DimensionMismatch("self-energy components must be contiguous complex128 fields of one shape")
    ctx = context(old[0].device.index or 0)
    out = ct.c_double()
    slot0 = ctx.result_slot() if ctx.deferring else 0
    ctx.call("qclhip_selfenergy_residual", ct.c_int64(n), (_vp * 3)(*[t.data_ptr() for t in old]),
             (_vp * 3)(*[t.data_ptr() for t in candidate]), ct.byref(out))
    return Deferred(slot0, 1) if ctx.deferring else out.value


# ---------------------------------------------------------------- Anderson
def _fields(ts, n=None):
    ts = list(ts)
    n = ts[0].numel() if n is None else n
    for t in ts:
        if t.dtype != torch.complex128 or t.numel() != n or not t.is_contiguous():
            raise _lib.ArgumentError("Anderson fields must be contiguous complex128 tensors of one size")
    return (_vp * len(ts))(*[t.data_ptr() for t in ts]), n


def anderson_append(state, cand, saved, resid, old_residuals):
    """Device part of _anderson_append! (production_mixing.jl:195-229):
    saved <- state, resid <- cand - state; returns dots (nfields, nold+1):
    Re <old_h, resid> per field and entry, then ||resid||^2 per field."""
    nf = len(state)
    ps, n = _fields(state)
    pc, _ = _fields(cand, n)
    psv, _ = _fields(saved, n)
    pr, _ = _fields(resid, n)
    nold = len(old_residuals)
    po = _fields([f for entry in old_residuals for f in entry], n)[0] if nold else None
    dots = np.zeros(nf * (nold + 1))
    ctx = context(state[0].device.index or 0)
    ctx.call("qclhip_anderson_append", ct.c_int64(n), nf, ps, pc, psv, pr, nold, po,
             dots.ctypes.data_as(_dp))
    return dots.reshape(nf, nold + 1)


def anderson_combine(coef, damping: float, X, R, out):
    """out_f = sum_h c_h (X[h][f] + damping R[h][f]) (production_mixing.jl:256-268)."""
    nf = len(out)
    po, n = _fields(out)
    px, _ = _fields([f for entry in X for f in entry], n)
    pr, _ = _fields([f for entry in R for f in entry], n)
    c = np.ascontiguousarray(coef, dtype=np.float64)
    ctx = context(out[0].device.index or 0)
    ctx.call("qclhip_anderson_combine", ct.c_int64(n), nf, len(c), c.ctypes.data_as(_dp), ct.c_double(damping),
             px, pr, po)


def anderson_guard(grid: Grid, families):
    """_anderson_guard (production_mixing.jl:82-111) over (R, L, G) device
    families: (first_failure one-based or 0, code)."""
    for fam in families:
        for t in fam:
            _check_field(grid, t, "self-energy")
    nf = len(families)
    arr = lambda k: (_vp * nf)(*[f[k].data_ptr() for f in families])
    first, code = ct.c_int64(), ct.c_int()
    grid.ctx.call("qclhip_anderson_guard", grid.handle, nf, arr(0), arr(1), arr(2), ct.byref(first),
                  ct.byref(code))
    return first.value, code.value


# ---------------------------------------------------------------- outer loop
def _dvec(x, device, dtype=torch.float64):
    if isinstance(x, torch.Tensor):
        if x.dtype != dtype or not x.is_contiguous() or x.device.type != "cuda":
            raise _lib.ArgumentError("expected a contiguous device tensor of the right dtype")
        return x
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device=device)


def basis_to_device(chi, device="cuda"):
    """Nz x Nb complex basis (numpy, any order) -> device column-major storage
    (torch shape (Nb, Nz))."""
    return torch.from_numpy(np.ascontiguousarray(np.asarray(chi, np.complex128).T)).to(device)


def electron_density(grid: Grid, GL, chi, gs: float):
    """_electron_density_bar (observables.jl:1-26): nbar on the position grid.
    chi: device basis from basis_to_device (Nb, Nz)."""
    _check_field(grid, GL, "G<")
    if chi.dtype != torch.complex128 or chi.dim() != 2 or chi.shape[0] != grid.Nb or not chi.is_contiguous():
        raise _lib.DimensionMismatch("density basis must be a contiguous (Nb, Nz) complex128 device tensor")
    Nz = chi.shape[1]
    nbar = torch.empty(Nz, dtype=torch.float64, device=grid.ctx.device)
    grid.ctx.call("qclhip_electron_density", grid.handle, _ptr(GL), Nz, _ptr(chi), ct.c_double(gs), _ptr(nbar))
    return nbar


def periodic_poisson(eps_r, wx, ND, lambda_P: float, nbar, U=None):
    """solve_periodic_poisson (poisson.jl:41-65) on the device: returns
    (U, zeta, r_P, neutrality), or (U, Deferred of 3) in deferred mode."""
    dev = nbar.device
    ctx = context(dev.index or 0)
    er, w, nd = (_dvec(x, dev) for x in (eps_r, wx, ND))
    N = nbar.numel()
    if any(t.numel() != N for t in (er, w, nd)):
        raise _lib.DimensionMismatch("Poisson inputs must have N_z entries")
    U = torch.empty(N, dtype=torch.float64, device=dev) if U is None else U
    out = (ct.c_double * 3)()
    slot0 = ctx.result_slot() if ctx.deferring else 0
    ctx.call("qclhip_periodic_poisson", N, _ptr(er), _ptr(w), _ptr(nd), ct.c_double(lambda_P), _ptr(nbar),
             _ptr(U), out)
    if ctx.deferring:
        return U, Deferred(slot0, 3)
    return U, out[0], out[1], out[2]


def project_hamiltonians(Phi, U, S, kin, Minv, h=None):
    """project_hamiltonians (basis.jl:288-308) with S = H0 - F (Z - z0) and
    kin_m = C kappa_m^2: device h (Nb, Nb, Nk). Phi from basis_to_device; S,
    Minv from matrix_to_device."""
    dev = U.device
    ctx = context(dev.index or 0)
    Nb, Nz = Phi.shape
    kin = _dvec(kin, dev)
    Nk = kin.numel()
    h = torch.empty((Nb, Nb, Nk), dtype=torch.complex128, device=dev) if h is None else h
    ctx.call("qclhip_project_hamiltonians", Nz, Nb, Nk, _ptr(Phi), _ptr(U), _ptr(S), _ptr(kin), _ptr(Minv), _ptr(h))
    return h


def mix_potential(U, cand, alpha: float):
    """U <- (1-alpha) U + alpha cand on the device; returns max |cand - U_old|."""
    ctx = context(U.device.index or 0)
    out = ct.c_double()
    slot0 = ctx.result_slot() if ctx.deferring else 0
    ctx.call("qclhip_mix_potential", U.numel(), _ptr(U), _ptr(cand), ct.c_double(alpha), ct.byref(out))
    return Deferred(slot0, 1) if ctx.deferring else out.value


def potential_gradient(hbar, Phi):
    """Host chain rule to a potential sampled on the real-space grid:
    dJ/dU_g = sum_m Re sum_ab conj(hbar_m[a,b]) conj(Phi[g,a]) Phi[g,b]
    (dh_m/dU_g = Phi[g,:]^* Phi[g,:]^T, basis.jl:300). hbar: (Nb,Nb,Nk) device
    tensor in the reference layout (element [b,a,m] is hbar_m[a,b]); Phi:
    (Ng, Nb) host array."""
    H = hbar.sum(dim=2).cpu().numpy().T  # sum over m; [a,b]
    P = np.asarray(Phi)
    return np.real(np.einsum("ab,ga,gb->g", np.conj(H), np.conj(P), P))

class Batch:
    """Independent instances with their own grids (uneven N_E, N_k, N_b),
    solved together: one launch per kernel and tile class (qclhip_batch_*)."""

    def __init__(self, grids):
        self.grids = list(grids)
        self.ctx = self.grids[0].ctx
        arr = (_vp * len(self.grids))(*[g.handle for g in self.grids])
        h = _vp()
        self.ctx.call("qclhip_batch_create", len(self.grids), arr, ct.byref(h))
        self.handle = h

    def __del__(self):
        try:
            if self.handle:
                self.ctx.lib.qclhip_batch_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def _batch_args(batch: Batch, hs, families, targets, states):
    n = len(batch.grids)
    nf = len(families[0])
    if any(len(f) != nf for f in families):
        raise _lib.ArgumentError("every instance needs the same number of families")
    for g, fams in zip(batch.grids, families):
        for fam in fams:
            for t in fam:
                _check_field(g, t, "self-energy")
    fam = lambda k: (_vp * (n * nf))(*[families[i][f][k].data_ptr() for i in range(n) for f in range(nf)])
    ptrs = lambda xs: (_vp * n)(*[x.data_ptr() for x in xs])
    tg = np.ascontiguousarray(targets, dtype=np.float64)
    return (ptrs(hs), nf, fam(0), fam(1), fam(2), tg.ctypes.data_as(_dp), ptrs([s.GR for s in states]),
            ptrs([s.GL for s in states]), ptrs([s.GG for s in states])), tg


def batch_set_mode(batch: Batch, mode: str):
    """'persistent' (one launch per kernel and tile class) or 'per_instance'."""
    modes = {"persistent": 0, "per_instance": 1}
    if mode not in modes:
        raise _lib.ArgumentError("batch mode must be 'persistent' or 'per_instance'")
    batch.ctx.call("qclhip_batch_set_mode", batch.handle, modes[mode])


def batch_mode(batch: Batch) -> str:
    return ("persistent", "per_instance")[int(batch.ctx.lib.qclhip_batch_get_mode(batch.handle))]


def batch_autotune(batch: Batch, hs, families, targets, states) -> dict:
    """qclhip_batch_autotune: times both launch modes on these arguments and
    keeps the faster; returns the times and the choice."""
    args, tg = _batch_args(batch, hs, families, targets, states)
    ms = (ct.c_double * 2)()
    batch.ctx.call("qclhip_batch_autotune", batch.handle, *args, ms)
    return dict(persistent_ms=ms[0], per_instance_ms=ms[1], mode=batch_mode(batch))


def batch_current_green(batch: Batch, hs, families, targets, states=None):
    """current_green_production for every instance of the batch. families[i]
    is instance i's list of (SR, SL, SG) triples (the same count for all).
    Returns (list of GreenState without A, list of info dicts), or (states,
    Deferred of 5 n values) in deferred-results mode."""
    if states is None:
        states = [GreenState(empty_field(g), empty_field(g), empty_field(g), None, None) for g in batch.grids]
    n = len(batch.grids)
    args, tg = _batch_args(batch, hs, families, targets, states)
    out = (ct.c_double * (5 * n))()
    slot0 = batch.ctx.result_slot() if batch.ctx.deferring else 0
    batch.ctx.call("qclhip_batch_current_green", batch.handle, *args, out)
    if batch.ctx.deferring:
        return states, Deferred(slot0, 5 * n)
    infos = [dict(lam=out[5 * i], number=out[5 * i + 1], holes=out[5 * i + 2], a=out[5 * i + 3], c=out[5 * i + 4])
             for i in range(n)]
    return states, infos
