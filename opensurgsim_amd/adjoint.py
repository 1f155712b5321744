"""Differentiable implicit-Euler rollouts on the GPU, for material identification.

`ImplicitEulerRollout` is a torch.autograd.Function over the C ABI. Its forward advances a tetrahedral
mesh N implicit steps from a given state with per-element Young's moduli E. When the whole state history
(2 (N + 1) ndof doubles) fits the rollout's history budget, the forward records every state in device
memory (fem_step_record) and the backward runs fem_adjoint_steps over it in one call, with no forward
step repeated. Otherwise the forward keeps the state every k steps (checkpoints) and the backward
recomputes each k-step segment from its checkpoint before running the adjoint steps through it. Either
way the backward returns dL/dE together with dL/dx_0 and dL/dv_0 of the initial state.

The forward map is the reference's implicit Euler step (Math/OdeSolverEulerImplicit.cpp:55-161, one
Newton iteration, Jacobi CG on the lower-triangle view) with the elements of
Physics/Fem3DElementTetrahedron.cpp and Physics/Fem3DElementCorotationalTetrahedron.cpp. The gradient
is exact for that map up to the CG tolerance; use a tight tolerance (1e-12) when comparing with finite
differences.
"""
import numpy as np
import torch

from . import femgpu

# PyTorch's HIP runtime has to start before the library's (system ROCm) one in a process: start it now,
# before a context of this module's user touches the GPU.
if torch.cuda.is_available():
    torch.cuda.init()


def _dev_zeros(n, device):
    return torch.zeros(n, dtype=torch.float64, device=device)


class Rollout:
    """A FemGpu context (mesh, boundary conditions, loads and options already set) bound to torch
    tensors on its device. The initial state is (x_init, v_init).

    checkpoint_every=None keeps the whole state history when it fits history_bytes of device memory
    (no recomputation in the backward) and checkpoints every 10 steps otherwise; an integer k always
    checkpoints every k steps."""

    def __init__(self, sim, x_init, v_init, steps, checkpoint_every=None, device="cuda", history_bytes=4 << 30):
        if steps < 1 or (checkpoint_every is not None and checkpoint_every < 1):
            raise ValueError("steps and checkpoint_every must be positive")
        self.sim = sim
        self.device = torch.device(device)
        self.n = sim.ndof
        self.ne = sim.num_elements
        self.steps = int(steps)
        self.full_history = checkpoint_every is None and 2 * (self.steps + 1) * self.n * 8 <= history_bytes
        self.k = self.steps if self.full_history else int(checkpoint_every or 10)
        self.history = None
        self.x_init = torch.as_tensor(np.asarray(x_init, dtype=np.float64).reshape(-1), device=self.device).clone()
        self.v_init = torch.as_tensor(np.asarray(v_init, dtype=np.float64).reshape(-1), device=self.device).clone()
        # order the context's work with torch's on the same stream
        sim.set_stream(torch.cuda.current_stream(self.device).cuda_stream)
        self.checkpoints = []
        self.final = None
        self.young = None

    def _get_state(self):
        x = torch.empty(self.n, dtype=torch.float64, device=self.device)
        v = torch.empty_like(x)
        self.sim.get_state_device(x.data_ptr(), v.data_ptr())
        return x, v

    def _set_state(self, x, v):
        self.sim.set_state_device(x.data_ptr(), v.data_ptr())

    def _advance(self, nsteps):
        rc = self.sim.step(nsteps)
        if rc != femgpu.FEM_OK:
            raise FloatingPointError("the rollout's state became non-finite")

    def forward(self, young):
        """Runs the rollout with moduli `young` (ne values); returns the final (x, v) as device tensors."""
        E = young.detach().to(torch.float64).cpu().numpy() if torch.is_tensor(young) else np.asarray(young)
        self.young = E.copy()
        self.sim.set_young_modulus(E)
        self._set_state(self.x_init, self.v_init)
        if self.full_history:
            if self.history is None:
                self.history = (torch.empty((self.steps + 1, self.n), dtype=torch.float64, device=self.device),
                                torch.empty((self.steps + 1, self.n), dtype=torch.float64, device=self.device))
            xh, vh = self.history
            if self.sim.step_record(self.steps, xh.data_ptr(), vh.data_ptr()) != femgpu.FEM_OK:
                raise FloatingPointError("the rollout's state became non-finite")
            self.final = (xh[self.steps].clone(), vh[self.steps].clone())
            return self.final[0].clone(), self.final[1].clone()
        self.checkpoints = [(self.x_init.clone(), self.v_init.clone())]
        done = 0
        while done < self.steps:
            n = min(self.k, self.steps - done)
            self._advance(n)
            done += n
            if done < self.steps:
                self.checkpoints.append(self._get_state())
        xf, vf = self._get_state()
        self.final = (xf.clone(), vf.clone())
        return xf, vf

    def backward(self, grad_x, grad_v):
        """dL/dE (ne), dL/dx_0, dL/dv_0 for the adjoints grad_x, grad_v of the final state."""
        if self.young is None:
            raise RuntimeError("forward() first")
        self.sim.set_young_modulus(self.young)
        lx = (grad_x.detach().to(torch.float64).clone() if grad_x is not None else _dev_zeros(self.n, self.device))
        lv = (grad_v.detach().to(torch.float64).clone() if grad_v is not None else _dev_zeros(self.n, self.device))
        gE = _dev_zeros(self.ne, self.device)
        if self.full_history:
            xh, vh = self.history
            self.sim.adjoint_steps(self.steps, xh.data_ptr(), vh.data_ptr(), lx.data_ptr(), lv.data_ptr(),
                                   gE.data_ptr())
            self.sim.synchronize()
            return gE, lx, lv
        for seg in range(len(self.checkpoints) - 1, -1, -1):
            start = seg * self.k
            length = min(self.k, self.steps - start)
            x, v = self.checkpoints[seg]
            self._set_state(x, v)
            # the segment's states (rows 0 .. length): recomputed from its checkpoint up to row length - 1; row
            # length is the next checkpoint (or the final state), which the same steps reproduce bit for bit
            xh = torch.empty((length + 1, self.n), dtype=torch.float64, device=self.device)
            vh = torch.empty_like(xh)
            rc = self.sim.step_record(length - 1, xh.data_ptr(), vh.data_ptr())
            if rc != femgpu.FEM_OK:
                raise FloatingPointError("the rollout's state became non-finite")
            xe, ve = self.checkpoints[seg + 1] if seg + 1 < len(self.checkpoints) else self.final
            xh[length].copy_(xe)
            vh[length].copy_(ve)
            self.sim.adjoint_steps(length, xh.data_ptr(), vh.data_ptr(), lx.data_ptr(), lv.data_ptr(), gE.data_ptr())
        self.sim.synchronize()
        return gE, lx, lv


class ImplicitEulerRollout(torch.autograd.Function):
    """(E, x_0, v_0) -> (x_N, v_N) through N implicit-Euler steps on the GPU."""

    @staticmethod
    def forward(ctx, young, x_init, v_init, rollout):
        rollout.x_init = x_init.detach().to(torch.float64).contiguous().clone()
        rollout.v_init = v_init.detach().to(torch.float64).contiguous().clone()
        x, v = rollout.forward(young)
        ctx.rollout = rollout
        ctx.young_dtype = young.dtype
        return x, v

    @staticmethod
    def backward(ctx, grad_x, grad_v):
        gE, gx0, gv0 = ctx.rollout.backward(grad_x, grad_v)
        return gE.to(ctx.young_dtype), gx0, gv0, None


def surface_loss(x_final, x_rest, obs_dofs, u_obs):
    """Squared displacement mismatch sum ||u - u_obs||^2 over the observed DOFs."""
    u = x_final - x_rest
    return torch.sum((u[obs_dofs] - u_obs) ** 2)


def recover_young_modulus(rollout, x_rest, obs_dofs, u_obs, young_init, iterations=10, callback=None):
    """Gradient-based identification of per-element moduli from observed displacements: L-BFGS with a
    strong-Wolfe line search on log E (every function evaluation is one forward + adjoint rollout).
    Returns (E, losses of the evaluations)."""
    dev = rollout.device
    logE = torch.log(torch.as_tensor(young_init, dtype=torch.float64, device=dev)).clone().requires_grad_(True)
    x_rest = torch.as_tensor(x_rest, dtype=torch.float64, device=dev).reshape(-1)
    obs = torch.as_tensor(obs_dofs, dtype=torch.long, device=dev)
    u_obs = torch.as_tensor(u_obs, dtype=torch.float64, device=dev).reshape(-1)
    opt = torch.optim.LBFGS([logE], lr=1.0, max_iter=iterations, history_size=10, tolerance_grad=1e-12,
                            tolerance_change=1e-14, line_search_fn="strong_wolfe")
    losses = []
    scale = []

    def closure():
        opt.zero_grad()
        x, _ = ImplicitEulerRollout.apply(torch.exp(logE), rollout.x_init, rollout.v_init, rollout)
        L = surface_loss(x, x_rest, obs, u_obs)
        if not scale:  # the optimizer works on L / L(first evaluation), an O(1) objective
            scale.append(max(float(L.detach()), 1e-300))
        (L / scale[0]).backward()
        losses.append(float(L.detach()))
        if callback:
            callback(len(losses) - 1, losses[-1])
        return L / scale[0]

    opt.step(closure)
    return torch.exp(logE.detach()), losses
