"""Learned SCF initial guess: an MLP maps local-geometry descriptors of an
atom pair to that pair's block of the alpha density matrix.

No reference counterpart: the reference starts from the core Hamiltonian
(Methods/hartree_fock.py:55-57).  The network is trained and evaluated with
the library's own FP32 kernels (pq_mlp_linear_fwd / _bwd, pq_mlp_mse,
pq_mlp_adam); torch only provides device buffers.

Density blocks are not rotation invariant, so each block is learned in a
local frame of its atom pair and rotated back:

* frame of an off-diagonal pair (A, B): e1 along B - A, e2 towards the atom
  nearest to A (other than A, B) orthogonalised against e1, e3 = e1 x e2;
  of a diagonal block (A, A): e1 towards the nearest neighbour, e2 towards
  the second nearest.
* a rotation R (rows e1, e2, e3) acts on the normalised Cartesian functions
  of a shell through M(R) (monomials of R v expanded in monomials of v,
  scaled by the per-component normalisation), and the density block in the
  local frame is D' = M_A^-T D M_B^-1, i.e. D = M_A^T D' M_B.
* descriptors (invariant): pair distance, 16 radial Gaussian features
  (centres 0.5..4 A, Z-weighted, cutoff 4 A) for each of the two atoms, and
  the local coordinates of the frame-defining neighbours.
* one network per (element A, element B, diagonal) type: 3 hidden layers of
  128 tanh units and a linear output of n_A * n_B values, inputs and outputs
  standardised.
"""
import ctypes
import itertools
import math
import time

import numpy as np
import torch

from . import _lib, structures
from ._lib import check

ANG = 1.0 / structures.TO_BOHR
RCUT_PAIR = 4.5          # Angstrom: pairs beyond this start from a zero block
RCUT_ENV = 4.0
MU = np.linspace(0.5, 4.0, 16)
ETA = 1.0 / (2 * 0.25 ** 2)
NFEAT = 1 + 16 + 16 + 4
HIDDEN = 128


# ----------------------------------------------------------------- geometry
def _monomial_expansion(l, R):
    """P[c][c'] with prod_i (sum_j R_ij v_j)^{c_i} = sum_c' P[c][c'] v^{c'}."""
    comps = structures.cart_components(l)
    index = {c: k for k, c in enumerate(comps)}
    P = np.zeros((len(comps), len(comps)))
    for a, c in enumerate(comps):
        # product over axes i of (R_i . v)^{c_i}: expand as a polynomial dict
        poly = {(0, 0, 0): 1.0}
        for i in range(3):
            for _ in range(c[i]):
                new = {}
                for mono, coef in poly.items():
                    for j in range(3):
                        m2 = list(mono)
                        m2[j] += 1
                        m2 = tuple(m2)
                        new[m2] = new.get(m2, 0.0) + coef * R[i, j]
                poly = new
        for mono, coef in poly.items():
            P[a, index[mono]] += coef
    return P


def cart_rotation(l, R):
    """M(R) acting on the normalised Cartesian functions of one shell."""
    if l == 0:
        return np.ones((1, 1))
    comps = structures.cart_components(l)
    N = np.array([(math.gamma(a + 0.5) * math.gamma(b + 0.5) * math.gamma(c + 0.5)) ** -0.5 for a, b, c in comps])
    P = _monomial_expansion(l, R)
    return (N[:, None] * P) / N[None, :]


def atom_rotation(ls, R):
    """Block-diagonal M(R) over an atom's shells (list of l)."""
    blocks = [cart_rotation(l, R) for l in ls]
    n = sum(b.shape[0] for b in blocks)
    M = np.zeros((n, n))
    o = 0
    for b in blocks:
        k = b.shape[0]
        M[o:o + k, o:o + k] = b
        o += k
    return M


def _unit(v):
    return v / np.linalg.norm(v)


def _frame(xyz, a, b):
    """Rows e1, e2, e3 of the local frame, and the frame-defining atoms."""
    d = np.linalg.norm(xyz - xyz[a], axis=1)
    order = [k for k in np.argsort(d) if k != a and k != b]
    if a != b:
        e1 = _unit(xyz[b] - xyz[a])
        refs = order
    else:
        if not order:
            return np.eye(3), []
        e1 = _unit(xyz[order[0]] - xyz[a])
        refs = order[1:]
    e2 = None
    used = []
    for k in refs:
        v = xyz[k] - xyz[a]
        w = v - np.dot(v, e1) * e1
        if np.linalg.norm(w) > 1e-3:
            e2 = _unit(w)
            used.append(k)
            break
    if e2 is None:
        t = np.array([1.0, 0.0, 0.0]) if abs(e1[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
        e2 = _unit(t - np.dot(t, e1) * e1)
    e3 = np.cross(e1, e2)
    if a == b:
        used = [order[0]] + used
    return np.array([e1, e2, e3]), used


def _radial(xyz, Z, a):
    d = np.linalg.norm(xyz - xyz[a], axis=1) * ANG
    m = (d > 1e-8) & (d < RCUT_ENV)
    fc = 0.5 * (np.cos(np.pi * d[m] / RCUT_ENV) + 1.0)
    return (Z[m, None] * fc[:, None] * np.exp(-ETA * (d[m, None] - MU[None, :]) ** 2)).sum(0)


def pair_samples(mol):
    """Yield (type, i, j, features, local->global rotation data) for all atom
    pairs i <= j within RCUT_PAIR (Angstrom)."""
    xyz = np.array([a.Coordinates for a in mol.Atoms])
    Z = np.array([a.NuclearCharge for a in mol.Atoms])
    rad = [_radial(xyz, Z, a) for a in range(len(Z))]
    out = []
    for i in range(len(Z)):
        for j in range(i, len(Z)):
            a, b = (i, j) if Z[i] >= Z[j] else (j, i)
            r = np.linalg.norm(xyz[b] - xyz[a]) * ANG
            if r > RCUT_PAIR:
                continue
            R, used = _frame(xyz, a, b)
            loc = [(R @ (xyz[k] - xyz[a])) * ANG for k in used[:2]]
            while len(loc) < 2:
                loc.append(np.zeros(3))
            f = np.concatenate([[r], rad[a], rad[b], [loc[0][0], loc[0][1], loc[1][0], loc[1][1]]])
            key = (mol.Atoms[a].Label, mol.Atoms[b].Label, a == b)
            out.append((key, a, b, f, R))
    return out


def _atom_functions(mol):
    """Per atom: function indices and shell l list (final = Cartesian basis)."""
    idx = [[] for _ in range(mol.NAtom)]
    ls = [[] for _ in range(mol.NAtom)]
    o = 0
    for at, atom in enumerate(mol.Atoms):
        for cg in atom.Basis:
            idx[at].extend(range(o, o + cg.NAngMom))
            ls[at].append(cg.AngularMomentum)
            o += cg.NAngMom
    return idx, ls


def local_block(D, mol, a, b, R, idx=None, ls=None):
    """D'_{ab} = M_a^-T D_ab M_b^-1 (block of D in the pair's local frame)."""
    if idx is None:
        idx, ls = _atom_functions(mol)
    Ma, Mb = atom_rotation(ls[a], R), atom_rotation(ls[b], R)
    blk = D[np.ix_(idx[a], idx[b])]
    return np.linalg.solve(Ma.T, blk) @ np.linalg.inv(Mb)


# ----------------------------------------------------------------- network
def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class MLP:
    """Input -> 3 x (128, tanh) -> linear output, FP32 on the GPU."""

    def __init__(self, nin, nout, seed=0, device="cuda"):
        g = np.random.default_rng(seed)
        dims = [nin, HIDDEN, HIDDEN, HIDDEN, nout]
        self.dims = dims
        self.W, self.b = [], []
        for i in range(4):
            lim = math.sqrt(6.0 / (dims[i] + dims[i + 1]))
            self.W.append(torch.tensor(g.uniform(-lim, lim, (dims[i], dims[i + 1])), dtype=torch.float32, device=device))
            self.b.append(torch.zeros(dims[i + 1], dtype=torch.float32, device=device))
        self.m = [torch.zeros_like(p) for p in self.W + self.b]
        self.v = [torch.zeros_like(p) for p in self.W + self.b]
        self.step = 0
        self.xmean = self.xstd = self.ymean = self.ystd = None

    def params(self):
        return self.W + self.b

    def forward(self, X, keep=False):
        L = _lib.lib()
        acts = [X]
        zs = []
        for i in range(4):
            B, I, O = X.shape[0], self.dims[i], self.dims[i + 1]
            Z = torch.empty((B, O), dtype=torch.float32, device=X.device)
            A = torch.empty_like(Z)
            act = 1 if i < 3 else 0
            check(L.pq_mlp_linear_fwd(_ptr(acts[-1]), _ptr(self.W[i]), _ptr(self.b[i]), _ptr(Z), _ptr(A),
                                      B, I, O, act, _stream()), "pq_mlp_linear_fwd")
            zs.append(Z)
            acts.append(A)
        return (acts[-1], acts) if keep else acts[-1]

    def backward(self, acts, dY):
        """Gradients of the parameters given dLoss/dY; returns (gW, gb)."""
        L = _lib.lib()
        gW = [torch.empty_like(w) for w in self.W]
        gb = [torch.empty_like(b) for b in self.b]
        dA = dY
        for i in reversed(range(4)):
            B, I, O = acts[i].shape[0], self.dims[i], self.dims[i + 1]
            dZ = torch.empty((B, O), dtype=torch.float32, device=dY.device)
            dX = torch.empty((B, I), dtype=torch.float32, device=dY.device) if i > 0 else None
            act = 1 if i < 3 else 0
            check(L.pq_mlp_linear_bwd(_ptr(acts[i]), _ptr(self.W[i]), _ptr(acts[i + 1]), _ptr(dA), _ptr(dZ),
                                      _ptr(dX) if dX is not None else None, _ptr(gW[i]), _ptr(gb[i]),
                                      B, I, O, act, _stream()), "pq_mlp_linear_bwd")
            dA = dX
        return gW, gb

    def loss_grad(self, Y, T):
        L = _lib.lib()
        dY = torch.empty_like(Y)
        loss = torch.zeros(1, dtype=torch.float32, device=Y.device)
        check(L.pq_mlp_mse(_ptr(Y), _ptr(T), None, _ptr(dY), _ptr(loss), Y.numel(), 1.0 / Y.numel(), _stream()),
              "pq_mlp_mse")
        return loss, dY

    def adam(self, gW, gb, lr=1e-3):
        L = _lib.lib()
        self.step += 1
        for p, g, m, v in zip(self.params(), gW + gb, self.m, self.v):
            check(L.pq_mlp_adam(_ptr(p), _ptr(g), _ptr(m), _ptr(v), p.numel(), lr, 0.9, 0.999, 1e-8, self.step,
                                _stream()), "pq_mlp_adam")

    def state(self):
        d = {"W%d" % i: w.cpu().numpy() for i, w in enumerate(self.W)}
        d.update({"b%d" % i: b.cpu().numpy() for i, b in enumerate(self.b)})
        d.update(xmean=self.xmean, xstd=self.xstd, ymean=self.ymean, ystd=self.ystd)
        return d

    @classmethod
    def from_state(cls, d, device="cuda"):
        W0 = d["W0"]
        net = cls(W0.shape[0], d["W3"].shape[1], device=device)
        for i in range(4):
            net.W[i] = torch.tensor(d["W%d" % i], device=device)
            net.b[i] = torch.tensor(d["b%d" % i], device=device)
        net.xmean, net.xstd, net.ymean, net.ystd = d["xmean"], d["xstd"], d["ymean"], d["ystd"]
        return net


# ----------------------------------------------------------------- data
def collect(mols, densities):
    """Training samples per pair type from converged alpha densities."""
    data = {}
    for mol, D in zip(mols, densities):
        idx, ls = _atom_functions(mol)
        for key, a, b, f, R in pair_samples(mol):
            blk = local_block(D, mol, a, b, R, idx, ls)
            data.setdefault(key, ([], []))
            data[key][0].append(f)
            data[key][1].append(blk.reshape(-1))
    return {k: (np.array(x, dtype=np.float32), np.array(y, dtype=np.float32)) for k, (x, y) in data.items()}


def train(data, epochs=200, batch=1024, lr=2e-3, seed=0, time_budget=None, log=None):
    """Train one network per pair type.  Returns {type: MLP}, {type: stats}."""
    nets, stats = {}, {}
    t0 = time.perf_counter()
    for ti, (key, (X, Y)) in enumerate(sorted(data.items())):
        net = MLP(X.shape[1], Y.shape[1], seed=seed + ti)
        net.xmean, net.xstd = X.mean(0), X.std(0) + 1e-6
        net.ymean, net.ystd = Y.mean(0), Y.std(0) + 1e-6
        Xd = torch.tensor((X - net.xmean) / net.xstd, device="cuda")
        Yd = torch.tensor((Y - net.ymean) / net.ystd, device="cuda")
        n = Xd.shape[0]
        g = torch.Generator(device="cpu").manual_seed(seed)
        first = last = None
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).to("cuda")
            tot = 0.0
            nb = 0
            for s in range(0, n, batch):
                ids = perm[s:s + batch]
                xb, yb = Xd[ids].contiguous(), Yd[ids].contiguous()
                out, acts = net.forward(xb, keep=True)
                loss, dY = net.loss_grad(out, yb)
                gW, gb = net.backward(acts, dY)
                net.adam(gW, gb, lr=lr * (0.5 ** (ep / max(1, epochs // 4))))
                tot += float(loss)
                nb += 1
            last = tot / nb
            first = last if first is None else first
            if time_budget and time.perf_counter() - t0 > time_budget * (ti + 1) / len(data):
                break
        nets[key] = net
        stats["%s-%s%s" % (key[0], key[1], "-diag" if key[2] else "")] = {"samples": int(n), "epochs": ep + 1,
                                                                          "loss_first": first, "loss_last": last}
        if log:
            log("type %s: %d samples, %d epochs, loss %.4g -> %.4g" % (key, n, ep + 1, first, last))
    return nets, stats


def predict_density(mol, nets):
    """Alpha-density guess assembled from the predicted local-frame blocks,
    renormalised to N_alpha electrons and McWeeny-purified twice."""
    idx, ls = _atom_functions(mol)
    n = mol.NOrbitals
    D = np.zeros((n, n))
    groups = {}
    for key, a, b, f, R in pair_samples(mol):
        groups.setdefault(key, []).append((a, b, f, R))
    for key, items in groups.items():
        net = nets.get(key)
        if net is None:
            continue
        X = np.array([f for _, _, f, _ in items], dtype=np.float32)
        Xd = torch.tensor((X - net.xmean) / net.xstd, device="cuda")
        Y = net.forward(Xd).cpu().numpy() * net.ystd + net.ymean
        for (a, b, f, R), y in zip(items, Y):
            Ma, Mb = atom_rotation(ls[a], R), atom_rotation(ls[b], R)
            blk = Ma.T @ y.reshape(len(idx[a]), len(idx[b])).astype(float) @ Mb
            D[np.ix_(idx[a], idx[b])] = blk
            if a != b:
                D[np.ix_(idx[b], idx[a])] = blk.T
            else:
                D[np.ix_(idx[a], idx[a])] = 0.5 * (blk + blk.T)
    return D


def purify(D, S, nalpha, rounds=2):
    """Scale to tr(DS) = N_alpha, then McWeeny D <- 3 DSD - 2 DSDSD."""
    D = D * (nalpha / max(1e-12, np.trace(D @ S)))
    for _ in range(rounds):
        DS = D @ S
        D = 3.0 * DS @ D - 2.0 * DS @ DS @ D
        D = 0.5 * (D + D.T)
    return D


def save(nets, path, meta=None):
    arrays = {}
    for key, net in nets.items():
        tag = "%s_%s_%d" % (key[0], key[1], int(key[2]))
        for k, v in net.state().items():
            arrays["%s__%s" % (tag, k)] = v
    if meta:
        arrays["__meta__"] = np.frombuffer(__import__("json").dumps(meta).encode(), dtype=np.uint8)
    np.savez(path, **arrays)


def load(path):
    z = np.load(path, allow_pickle=False)
    groups = {}
    for name in z.files:
        if name == "__meta__":
            continue
        tag, k = name.split("__")
        groups.setdefault(tag, {})[k] = z[name]
    nets = {}
    for tag, d in groups.items():
        a, b, diag = tag.split("_")
        nets[(a, b, bool(int(diag)))] = MLP.from_state(d)
    return nets


def training_molecules(count, seed=5, sigma=0.05):
    """Jittered C1..C8 alkanes and primary alcohols, 6-31G* (Cartesian d)."""
    from . import geometry
    rng = np.random.default_rng(seed)
    mols = []
    for k in range(count):
        n = int(rng.integers(1, 9))
        rows = geometry.alkane(n, sigma, int(rng.integers(1 << 30))) if rng.random() < 0.5 else \
            geometry.alcohol(n, sigma, int(rng.integers(1 << 30)))
        mols.append(structures.Molecule(rows, basis="6-31G*", cartesian_l=[2]))
    return mols
