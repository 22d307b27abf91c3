"""sfsmodel/sfsmodel_gradient.py:39-148: the gradient closure ('tensor'
model).  update() is one HIP kernel (pf_sfs_gradient) that writes the six
padded stress fields from the cell-centred velocity gradients."""

import torch


class residual_stress:
    def __init__(self, decomp, geo, metric):
        prec = decomp.prec
        self.modelType = 'tensor'
        self.box = decomp.box
        shape = (decomp.nxo_, decomp.nyo_, decomp.nzo_)
        for n in ('FXX', 'FYY', 'FZZ', 'FXY', 'FYZ', 'FXZ'):
            setattr(self, n, torch.zeros(shape, dtype=prec, device=decomp.device))
        if geo.type == 'uniform':
            self.const = geo.dx * geo.dx / 12.0
        else:
            raise Exception('sfsmodel_gradient: geometry type not implemented')

    def update(self, state_u, state_v, state_w, metric=None):
        self.box.sfs_gradient(state_u.var, state_v.var, state_w.var,
                              (self.FXX, self.FYY, self.FZZ, self.FXY, self.FXZ, self.FYZ))
