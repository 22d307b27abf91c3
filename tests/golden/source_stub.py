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
