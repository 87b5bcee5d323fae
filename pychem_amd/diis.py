"""Pulay DIIS on the SDF - FDS residual in the orthogonal basis, FP64 on the
host (Methods/hf_extensions/diis.py:14-155).

Behaviour kept from the reference: a residual enters the subspace only while
its max element is below 0.5 (diis.py:22,33); the oldest vectors are dropped
while the B-matrix condition number exceeds DIIS_Max_Condition (1e7,
diis.py:87-97; the reference's size clause never fires, diis.py:90); C1
coefficients from solving the bordered B system (diis.py:108-112).
"""
import numpy as np

# Works on numpy arrays or torch tensors: only @, *, sum and indexing are used
# on the N x N matrices; the small B matrix is numpy.


class DIIS:
    def __init__(self, threshold=0.5, max_condition=1e7, max_size=None):
        self.threshold = threshold
        self.max_condition = max_condition
        self.max_size = max_size
        self.focks = []
        self.residuals = []
        self.matrix = None
        self.error = 1.0

    @staticmethod
    def residual(S, D, F, X):
        """Xt (S D F - F D S) X, zero-padded to N x N (diis.py:48-54); with
        S, D, F symmetric, F D S is the transpose of S D F."""
        A = S @ D @ F
        r = A - A.T
        t = X.T @ r @ X
        out = r * 0.0
        n = t.shape[0]
        out[:n, :n] = t
        return out

    @staticmethod
    def _dot(a, b):
        return float((a * b).sum())

    @staticmethod
    def _dots(a, bs):
        """<a, b> for every b, with one device-to-host transfer for tensors."""
        if hasattr(a, "device") and getattr(a.device, "type", "cpu") != "cpu":
            import torch
            return torch.stack([(a * b).sum() for b in bs]).tolist()
        return [float((a * b).sum()) for b in bs]

    @staticmethod
    def _cond(m):
        ev = np.abs(np.linalg.eigvals(m))
        return np.inf if ev.min() == 0.0 else ev.max() / ev.min()

    def extrapolate(self, res, F):
        if hasattr(res, "device") and getattr(res.device, "type", "cpu") != "cpu":
            import torch
            self.error, amax = torch.stack([res.max(), res.abs().max()]).tolist()
        else:
            self.error, amax = float(res.max()), float(abs(res).max())
        if abs(self.error) >= abs(self.threshold):
            return F
        if amax < 1e-14:
            return F          # converged: nothing to extrapolate
        self.residuals.append(res)
        self.focks.append(F)
        if len(self.residuals) < 2:
            return F
        size = len(self.residuals) + 1
        B = np.full((size, size), -1.0)
        B[-1, -1] = 0.0
        if self.matrix is None:
            for i, r1 in enumerate(self.residuals):
                B[i, :size - 1] = self._dots(r1, self.residuals)
        else:
            B[:size - 1, :size - 1] = self.matrix
            new = self._dots(self.residuals[-1], self.residuals)
            B[:len(new), size - 2] = new
            B[size - 2, :len(new)] = new
        self.matrix = B
        while (self.max_size is not None and len(self.residuals) > self.max_size) or \
                self._cond(self.matrix) > self.max_condition:
            if len(self.residuals) == 1:
                break
            self.matrix = np.delete(np.delete(self.matrix, 0, 0), 0, 1)
            self.residuals.pop(0)
            self.focks.pop(0)
        rhs = np.zeros(len(self.matrix))
        rhs[-1] = -1.0
        try:
            coeffs = np.linalg.solve(self.matrix, rhs)[:-1]
        except np.linalg.LinAlgError:
            coeffs = np.linalg.lstsq(self.matrix, rhs, rcond=None)[0][:-1]
        out = F * 0.0
        for c, f in zip(coeffs, self.focks):
            out += float(c) * f
        return out
