"""torch.autograd.Function over the engine's differentiable rollout (test
harness).  Forward runs sph_rollout_forward from (pos0, v0) with per-scene
stiffness / viscosity; backward calls sph_rollout_backward, which recomputes
checkpointed segments and runs the hand-written adjoint kernels."""
import numpy as np
import torch

from simplefluid_amd import engine as E


class SPHRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, stiffness, viscosity, eng, base_params, pos0, steps, every):
        n = pos0.shape[0]
        for s in range(eng.num_scenes):
            p = E.clone_params(base_params[s], pressure_stiffness=float(stiffness[s]),
                               viscosity=float(viscosity[s]))
            if isinstance(eng, E.BatchEngine):
                eng.set_scene_params(s, p)
            else:
                eng.set_params(p)
        eng.set_particles(np.asarray(pos0, np.float32), v0.detach().cpu().numpy().astype(np.float32))
        eng.make_ready()
        eng.rollout_forward(int(steps), int(every))
        ctx.eng = eng
        xT = eng.get_particles()
        return torch.from_numpy(xT.astype(np.float64)).reshape(n, 3)

    @staticmethod
    def backward(ctx, gx):
        dB, dmu, dv0 = ctx.eng.rollout_backward(gx.detach().cpu().numpy().astype(np.float32))
        return (torch.from_numpy(dv0.astype(np.float64)), torch.from_numpy(dB.astype(np.float64)),
                torch.from_numpy(dmu.astype(np.float64)), None, None, None, None, None)


def rollout(v0, stiffness, viscosity, eng, base_params, pos0, steps, every=20):
    return SPHRollout.apply(v0, stiffness, viscosity, eng, base_params, pos0, steps, every)
