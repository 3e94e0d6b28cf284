"""Rate-and-state (fast-velocity-weakening) friction solver on the GPU.

Mirrors ``seissol::dr::friction_law::FrictionSolver`` for friction law 103
(src/DynamicRupture/FrictionLaws/FrictionSolver.h:28-108):
``setupLayer``, ``evaluate(fullUpdateTime, frictionTime, timeWeights)``,
``computeDeltaT``, ``allocationPlace``, ``clone``; all arithmetic in the HIP
library (``sdg_friction_evaluate``).

Also the synthetic planar-fault workload (one layer of faces with the Stroud
Gauss points of each face) and the warm-start MLP trainer, whose training and
inference kernels are the library's (``sdg_mlp_*``).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _capi
from ._capi import FAULT_FIELDS, FaultLayer, FrictionOptions, FrictionParams, check, lib, ptr

MPA = 1e6


def dr_quadrature(order: int):
    n = (order + 1) ** 2
    pts, w, R = np.zeros((n, 2)), np.zeros(n), np.zeros((n, n))
    check(lib().sdg_dr_quadrature(order, pts.ctypes.data, w.ctypes.data, R.ctypes.data), "sdg_dr_quadrature")
    return pts, w, R


def gauss_legendre(n: int, t0: float, t1: float):
    p, w = np.zeros(n), np.zeros(n)
    check(lib().sdg_gauss_legendre(n, t0, t1, p.ctypes.data, w.ctypes.data), "sdg_gauss_legendre")
    return p, w


def compute_delta_t(time_points) -> np.ndarray:
    tp = np.ascontiguousarray(time_points, dtype=np.float64)
    out = np.zeros_like(tp)
    check(lib().sdg_friction_delta_t(len(tp), tp.ctypes.data, out.ctypes.data), "sdg_friction_delta_t")
    return out


@dataclass
class FrictionParameters:
    """DRParameters subset (src/Initializer/Parameters/DRParameters.h:55-76)."""

    rs_sr0: float = 1e-6
    rs_slip_rate_tolerance: float = 1e-8
    rs_state_tolerance: float = 1e-8
    rs_max_slip_rate_updates: int = 60
    rs_state_variable_updates: int = 10
    eta_damp: float = 1.0
    eta_damp_end: float = math.inf
    nucleation_count: int = 1
    t0: float = 0.02
    s0: float = 0.0

    def c_struct(self) -> FrictionParams:
        return FrictionParams(self.rs_sr0, self.rs_slip_rate_tolerance, self.rs_state_tolerance,
                              self.rs_max_slip_rate_updates, self.rs_state_variable_updates, self.eta_damp,
                              self.eta_damp_end, self.nucleation_count, 0, self.t0, self.s0)


@dataclass
class FaultSetup:
    """Host FP64 state of one layer of fault faces ([faces][points] arrays)."""

    order: int
    num_faces: int
    points: int
    num_points: int
    arrays: dict
    positions: np.ndarray  # [faces][points][2] in-plane coordinates (m)
    params: FrictionParameters
    pulse: dict = field(default_factory=dict)


def make_fault_setup(num_faces: int = 20_000, order: int = 4, seed: int = 7, points: int | None = None,
                     params: FrictionParameters | None = None, length: float = 10_000.0,
                     nucleation_radius: float = 1_500.0, nucleation_stress: float = 60 * MPA,
                     rho: float = 2670.0, vp: float = 6000.0, vs: float = 3464.0, f0: float = 0.6,
                     mu_w: float = 0.2, v_w: float = 0.1, v_ini: float = 1e-16) -> FaultSetup:
    """Planar vertical fault: `num_faces` triangles of a square grid over a
    `length` x `length` plane, Stroud points of order `order` on each face.
    a ~ U[0.008, 0.02], b = 0.014, L ~ U[0.2, 0.5] m, sigma_n = -120 MPa,
    initial shear U[60, 80] MPa, Gaussian nucleation patch at the centre;
    f0, V0, Vw, mu_w and V_ini as in the SCEC TPV104 benchmark."""
    params = params or FrictionParameters()
    rng = np.random.default_rng(seed)
    qp, _, _ = dr_quadrature(order)
    npnt = qp.shape[0]
    P = points or npnt
    nsq = max(1, int(round(math.sqrt(num_faces / 2))))
    h = length / nsq
    pos = np.zeros((num_faces, P, 2))
    for f in range(num_faces):
        sq, upper = divmod(f, 2)
        i, j = divmod(sq % (nsq * nsq), nsq)
        x0, y0 = i * h, j * h
        v = (np.array([[x0, y0], [x0 + h, y0], [x0, y0 + h]]) if upper == 0
             else np.array([[x0 + h, y0 + h], [x0, y0 + h], [x0 + h, y0]]))
        pos[f, :npnt] = v[0] + np.outer(qp[:, 0], v[1] - v[0]) + np.outer(qp[:, 1], v[2] - v[0])
        pos[f, npnt:] = pos[f, 0]
    shape = (num_faces, P)
    A = {}
    A["rs_a"] = rng.uniform(0.008, 0.02, shape)
    A["rs_b"] = np.full(shape, 0.014)
    A["rs_sl0"] = rng.uniform(0.2, 0.5, shape)
    A["rs_f0"] = np.full(shape, f0)
    A["rs_muw"] = np.full(shape, mu_w)
    A["rs_srw"] = np.full(shape, v_w)
    ist = np.zeros((num_faces, 6, P))
    ist[:, 0, :] = -120 * MPA
    ist[:, 3, :] = rng.uniform(60 * MPA, 80 * MPA, shape)
    A["initial_stress"] = ist
    A["initial_pressure"] = np.zeros(shape)
    centre = np.array([length / 2, length / 2])
    r2 = ((pos - centre) ** 2).sum(-1)
    nuc = np.zeros((num_faces, 6, P))
    nuc[:, 3, :] = nucleation_stress * np.exp(-r2 / nucleation_radius ** 2)
    A["nucleation_stress"] = nuc
    A["nucleation_pressure"] = np.zeros(shape)
    # initial state: steady sliding at V_ini under the initial tractions
    sv = np.zeros(shape)
    mu = np.zeros(shape)
    L = lib()
    s_ = C.c_double()
    m_ = C.c_double()
    for f in range(num_faces):
        for p in range(P):
            check(L.sdg_rs_initial_state(ist[f, 3, p], ist[f, 5, p], ist[f, 0, p], A["rs_a"][f, p], params.rs_sr0,
                                         v_ini, C.byref(s_), C.byref(m_)), "sdg_rs_initial_state")
            sv[f, p], mu[f, p] = s_.value, m_.value
    A["state_variable"] = sv
    A["mu"] = mu
    A["slip_rate1"] = np.full(shape, v_ini)
    for k in ("slip_rate2", "slip_rate_magnitude", "traction1", "traction2", "slip1", "slip2", "accumulated_slip",
              "rupture_time", "peak_slip_rate", "dyn_stress_time"):
        A[k] = np.zeros(shape)
    for k in ("rupture_time_pending", "dyn_stress_time_pending", "convergence_inner", "convergence_outer"):
        A[k] = np.ones(shape, dtype=np.uint8)
    A["newton_iterations"] = np.zeros(shape, dtype=np.int32)
    A["slip_rate_updates"] = np.zeros(shape, dtype=np.int32)
    zp, zs = rho * vp, rho * vs
    ie = np.array([zp, zs, zp, zs, 1 / (1 / zp + 1 / zp), 1 / (1 / zs + 1 / zs), 1 / zs + 1 / zs, 1 / zp, 1 / zs,
                   1 / zp, 1 / zs])
    A["imp_and_eta"] = np.tile(ie, (num_faces, 1))
    A["q_interpolated_plus"] = np.zeros((num_faces, order, 9, P))
    A["q_interpolated_minus"] = np.zeros((num_faces, order, 9, P))
    A["imposed_state_plus"] = np.zeros((num_faces, 9, P))
    A["imposed_state_minus"] = np.zeros((num_faces, 9, P))
    return FaultSetup(order, num_faces, P, npnt, A, pos, params,
                      pulse={"centre": centre, "speed": 3000.0, "width": 400.0, "amplitude": 50 * MPA})


def pulse_loading(setup: FaultSetup, t_points: np.ndarray) -> np.ndarray:
    """Shear-stress perturbation of an expanding ring front (the same on both
    sides of the fault), [faces][order][points]: a synthetic stand-in for the
    waves that the coupled element update would bring to the fault."""
    pl = setup.pulse
    r = np.sqrt(((setup.positions - pl["centre"]) ** 2).sum(-1))  # [f][p]
    out = np.empty((setup.num_faces, len(t_points), setup.points))
    for o, t in enumerate(t_points):
        out[:, o, :] = pl["amplitude"] * np.exp(-((r - pl["speed"] * t) / pl["width"]) ** 2)
    return out


class DeviceFault:
    """Device copies of a FaultSetup and the ``sdg_fault_layer`` view."""

    def __init__(self, setup: FaultSetup, device: str = "cuda"):
        import torch
        self.setup = setup
        self.t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(device) for k, v in setup.arrays.items()}
        self.view = FaultLayer(setup.num_faces, setup.points, setup.num_points, setup.order, 0,
                               *[ptr(self.t.get(name)) for name in FAULT_FIELDS])

    def host(self, name: str) -> np.ndarray:
        return self.t[name].cpu().numpy()


class FrictionSolver:
    """Device friction solver for law 103 (FrictionSolver.h:38-65).

    Warm start of the first slip-rate solve of each time point:
    ``closed_form_start`` starts creeping points at the closed-form creep root,
    ``mlp_params`` (device FP32, ``mlp_param_count()`` values) starts the other
    points at the network's predicted root."""

    def __init__(self, params: FrictionParameters, order: int, resample_order: int | None = None):
        import torch
        self.params = params
        self.order = order
        _, _, R = dr_quadrature(resample_order or order)
        self.resample = torch.from_numpy(R).cuda()
        self.layer = None
        self.mlp_params = None
        self.closed_form_start = False

    @staticmethod
    def computeDeltaT(time_points):
        return compute_delta_t(time_points)

    @staticmethod
    def allocationPlace():
        return "device"

    def clone(self):
        other = FrictionSolver.__new__(FrictionSolver)
        other.__dict__.update(self.__dict__)
        return other

    def setupLayer(self, layer: DeviceFault):
        self.layer = layer

    def evaluate(self, fullUpdateTime: float, deltaT, timeWeights, options: FrictionOptions | None = None,
                 stream=None):
        import torch
        dt = np.ascontiguousarray(deltaT, dtype=np.float64)
        tw = np.ascontiguousarray(timeWeights, dtype=np.float64)
        opt = FrictionOptions() if options is None else FrictionOptions.from_buffer_copy(options)
        if self.mlp_params is not None and not opt.mlp_params:
            opt.mlp_params = ptr(self.mlp_params)
        if self.closed_form_start:
            opt.closed_form_start = 1
        s = torch.cuda.current_stream().cuda_stream if stream is None else stream
        prm = self.params.c_struct()
        check(lib().sdg_friction_evaluate(C.byref(prm), C.byref(self.layer.view), float(fullUpdateTime),
                                          dt.ctypes.data, tw.ctypes.data, ptr(self.resample), C.byref(opt), s),
              "sdg_friction_evaluate")


def mlp_param_count() -> int:
    return int(lib().sdg_mlp_param_count())


class WarmStartTrainer:
    """Trains the 6-32-32-1 warm-start network with the library's kernels
    (loss/gradient, Adam) on samples recorded from the plain Newton solver."""

    def __init__(self, seed: int = 0):
        import torch
        n = mlp_param_count()
        rng = np.random.default_rng(seed)
        p = np.zeros(n, dtype=np.float32)
        # normal initialisation scaled by 1/sqrt(fan-in); zero biases
        p[14:14 + 192] = rng.normal(0, 1 / math.sqrt(6), 192)              # W1
        p[14 + 192 + 32:14 + 192 + 32 + 1024] = rng.normal(0, 1 / math.sqrt(32), 1024)  # W2
        p[14 + 192 + 32 + 1024 + 32:14 + 192 + 32 + 1024 + 64] = rng.normal(0, 1 / math.sqrt(32), 32)  # W3
        self.params = torch.from_numpy(p).cuda()
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grads = torch.zeros_like(self.params)
        self.loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        self.step = 0

    def set_normalisation(self, samples):
        x = samples[:, :6]
        y = samples[:, 6]
        self.params[0:6] = x.mean(0)
        self.params[6:12] = x.std(0).clamp_min(1e-3)
        self.params[12] = y.mean()
        self.params[13] = y.std().clamp_min(1e-3)

    def loss_grad(self, batch) -> float:
        check(lib().sdg_mlp_loss_grad(ptr(self.params), ptr(batch), batch.shape[0], ptr(self.grads), ptr(self.loss),
                                      None), "sdg_mlp_loss_grad")
        return self.loss

    def adam(self, lr: float):
        self.step += 1
        check(lib().sdg_mlp_adam(ptr(self.params), ptr(self.grads), ptr(self.m), ptr(self.v), self.step, lr, None),
              "sdg_mlp_adam")

    def fit(self, samples, steps: int = 3000, batch: int = 1024, lr: float = 3e-3, seed: int = 0,
            balance: bool = True) -> list:
        """`steps` Adam steps on random batches, learning rate decaying to lr/20.

        With `balance`, the solves whose target is non-negligible (|target| >
        0.01, the closed-form estimate is off) are mixed 1:1 with the others,
        which would otherwise dominate the loss with near-zero targets."""
        import torch
        g = torch.Generator(device="cuda").manual_seed(seed)
        if balance:
            off = samples[:, 6].abs() > 0.01
            hard, easy = samples[off], samples[~off]
            if 0 < hard.shape[0] < easy.shape[0]:
                pick = torch.randint(0, easy.shape[0], (hard.shape[0],), generator=g, device="cuda")
                samples = torch.cat([hard, easy[pick]]).contiguous()
        self.set_normalisation(samples)
        n = samples.shape[0]
        history = []
        for k in range(steps):
            idx = torch.randint(0, n, (min(batch, n),), generator=g, device="cuda")
            self.loss_grad(samples[idx].contiguous())
            self.adam(lr * (0.05 ** (k / max(1, steps - 1))))
            if k % 500 == 0 or k == steps - 1:
                history.append(float(self.loss.item()))
        torch.cuda.synchronize()
        return history

    def predict(self, features):
        import torch
        out = torch.zeros(features.shape[0], dtype=torch.float32, device="cuda")
        check(lib().sdg_mlp_forward(ptr(self.params), ptr(features), features.shape[0], features.shape[1], ptr(out),
                                    None), "sdg_mlp_forward")
        return out


def step_times(order: int, dt: float):
    """Gauss-Legendre time points/weights of a DR step and the friction
    sub-step widths (TimeCluster.cpp:301, FrictionSolver.cpp:20-35)."""
    tp, tw = gauss_legendre(order, 0.0, dt)
    return tp, tw, compute_delta_t(tp)


class FaultDriver:
    """Advances one device fault layer through `steps` DR steps of width `dt`
    under the synthetic ring-front loading (imposed through the interpolated
    shear stress of both sides)."""

    def __init__(self, setup: FaultSetup, solver: FrictionSolver, dt: float = 1e-3):
        import torch
        self.setup, self.solver, self.dt = setup, solver, dt
        self.fault = DeviceFault(setup)
        solver.setupLayer(self.fault)
        pl = setup.pulse
        pos = torch.from_numpy(setup.positions).cuda()
        self.r = torch.sqrt(((pos - torch.from_numpy(pl["centre"]).cuda()) ** 2).sum(-1))  # [f][p]
        self.tp, self.tw, self.delta = step_times(setup.order, dt)
        self.step_index = 0

    def load(self, t0: float):
        pl = self.setup.pulse
        for o, t in enumerate(self.tp):
            s = pl["amplitude"] * torch_exp(-((self.r - pl["speed"] * (t0 + t)) / pl["width"]) ** 2)
            self.fault.t["q_interpolated_plus"][:, o, 3, :] = s
            self.fault.t["q_interpolated_minus"][:, o, 3, :] = s

    def advance(self, steps: int, options: FrictionOptions | None = None):
        for _ in range(steps):
            t0 = self.step_index * self.dt
            self.load(t0)
            self.solver.evaluate(t0, self.delta, self.tw, options)
            self.step_index += 1


def torch_exp(x):
    import torch
    return torch.exp(x)
