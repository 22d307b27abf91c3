"""A deterministic stand-in for a 'source' closure (modelType 'source', the
contract of velocity.py:360-383 and sfsmodel_ML.py:354-404): update() keeps
copies of the three arrays it receives and sets GX, GY, GZ to fixed
elementwise polynomials of them, in their dtype and on their device, with the
(nxo, nyo, nzo, 1) shape the reference's closure gives them
(sfsmodel_ML.py:146-151).  The golden vectors record the reference's RHS and
RK4/Euler steps with it; the GPU tests run the same stand-in, so the inputs
the closure receives and the RHS are compared bit for bit.
"""


class RecordingSource:
    modelType = 'source'

    def __init__(self):
        self.calls = []

    def update(self, ui, vi, wi, requires_grad=False):
        self.calls.append((ui.clone(), vi.clone(), wi.clone()))
        self.GX = (0.3 * ui * vi - 0.2 * wi).unsqueeze(3)
        self.GY = (0.25 * vi * wi + 0.1 * ui).unsqueeze(3)
        self.GZ = (0.2 * wi * ui - 0.15 * vi).unsqueeze(3)


# Parameters of the reference's NeuralNetworkModel2 (sfsmodel_ML.py:76-87)
# for the golden vectors and tests: seeded draws in torch.nn.Linear's default
# ranges (uniform in +-1/sqrt(fan_in)), packed in the order and layout the
# HIP closure takes (weight[out][in] then bias, for fc1, fcG, fcG2, fc2, fc3,
# fc4).
ML_LAYERS = (('fc1', 200, 273), ('fcG', 200, 273), ('fcG2', 200, 273),
             ('fc2', 200, 200), ('fc3', 200, 200), ('fc4', 18, 200))


def ml_source_params(seed=7):
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    parts = []
    for _, n_out, n_in in ML_LAYERS:
        bound = 1.0 / np.sqrt(n_in)
        parts.append(rng.uniform(-bound, bound, size=n_out * n_in).astype(np.float32))
        parts.append(rng.uniform(-bound, bound, size=n_out).astype(np.float32))
    return np.concatenate(parts)


def ml_state_dict(flat):
    """Packed vector -> the reference model's state_dict (float32 CPU tensors)."""
    import torch
    flat = torch.as_tensor(flat, dtype=torch.float32).cpu()
    sd, o = {}, 0
    for name, n_out, n_in in ML_LAYERS:
        sd[name + '.weight'] = flat[o:o + n_out * n_in].reshape(n_out, n_in).clone()
        o += n_out * n_in
        sd[name + '.bias'] = flat[o:o + n_out].clone()
        o += n_out
    return sd
