"""Finite-element model generators that produce SAM control arrays.

The SPR solver consumes a model through SAM's control matrices (all 1-based,
INTEGER*8):

* MPAR(50)   -- MPAR(1)=NANOD, MPAR(2)=NEL, MPAR(3)=NDOF, MPAR(7)=NCEQ
  (sprprp.f:47-53, sprseq.f:31, sprsas.f:147-157)
* MADOF(NANOD+1) -- accumulated dofs per node
* MSC(NDOF)  -- status code per dof: 1 free, 0 fixed (sprseq.f:61-78)
* MPMNPC(NEL+1), MMNPC -- element -> node connectivity
* MPMCEQ(NCEQ+1), MMCEQ, TTCC -- constraint equations: equation I has its
  constrained dof MMCEQ(MPMCEQ(I)) (status code 0) with TTCC(MPMCEQ(I)) = c0,
  and masters MMCEQ(MPMCEQ(I)+1 .. MPMCEQ(I+1)-1) with their coefficients, so
  u_slave = c0 + sum(TTCC * u_master); no masters: a prescribed dof
  (sprseq.f:108-140)
* MINEX(NANOD) -- external node numbers (used in messages only)

Each generator also returns the element matrices, column-major, in the element
dof order that SPRADM expects (nodes in MMNPC order, each node's dofs in MADOF
order; spradm.f:88-104), and the random load vectors.  Random draws use
``numpy.random.default_rng(seed)`` (PCG64) in natural dof order (node-major,
dof-minor) before any equation renumbering.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

I8 = np.int64


@dataclass
class SamModel:
    """A finite-element model in SAM form plus its element matrices."""

    name: str
    nanod: int
    nel: int
    ndof: int
    madof: np.ndarray          # (nanod+1,) int64, 1-based
    msc: np.ndarray            # (ndof,) int64
    mpmnpc: np.ndarray         # (nel+1,) int64, 1-based
    mmnpc: np.ndarray          # (nnpc,) int64, 1-based node numbers
    nedof: np.ndarray          # (nel,) int64 element dof counts
    em: np.ndarray             # concatenated column-major element matrices (float64)
    em_off: np.ndarray         # (nel,) int64 offsets into em
    escale: np.ndarray | None  # optional per-element multiplier on em
    meta: dict = field(default_factory=dict)
    mpmceq: np.ndarray | None = None  # (nceq+1,) int64, 1-based; None: no constraints
    mmceq: np.ndarray | None = None   # (nmmceq,) int64, 1-based dofs
    ttcc: np.ndarray | None = None    # (nmmceq,) float64

    @property
    def nceq(self) -> int:
        return 0 if self.mpmceq is None else int(self.mpmceq.size - 1)

    def mpar(self) -> np.ndarray:
        mpar = np.zeros(50, dtype=I8)
        mpar[0] = self.nanod
        mpar[1] = self.nel
        mpar[2] = self.ndof
        mpar[6] = self.nceq  # NCEQ
        return mpar

    def ceq_arrays(self):
        """(MPMCEQ, MMCEQ, TTCC) as the reference routines take them; one-entry
        placeholders when the model has no constraint equations."""
        if self.mpmceq is None:
            return np.ones(1, dtype=I8), np.zeros(1, dtype=I8), np.zeros(1)
        return (np.ascontiguousarray(self.mpmceq, dtype=I8),
                np.ascontiguousarray(self.mmceq, dtype=I8),
                np.ascontiguousarray(self.ttcc, dtype=np.float64))

    def element_matrix(self, iel: int) -> np.ndarray:
        """Element matrix of element ``iel`` (0-based) as a dense array."""
        n = int(self.nedof[iel])
        k = self.em[self.em_off[iel]:self.em_off[iel] + n * n].reshape(n, n, order="F")
        if self.escale is not None:
            k = k * self.escale[iel]
        return k

    def element_dofs(self, iel: int) -> np.ndarray:
        """0-based global dof numbers of element ``iel`` in element-matrix order."""
        nodes = self.mmnpc[self.mpmnpc[iel] - 1:self.mpmnpc[iel + 1] - 1] - 1
        return np.concatenate([np.arange(self.madof[n] - 1, self.madof[n + 1] - 1) for n in nodes])


# --------------------------------------------------------------------------
# Element matrices
# --------------------------------------------------------------------------

_GAUSS2 = np.array([-1.0, 1.0]) / np.sqrt(3.0)


def hex8_stiffness(E: float = 1.0, nu: float = 0.3, h: float = 1.0) -> np.ndarray:
    """24x24 stiffness of a trilinear hexahedron (cube of edge h), 2x2x2 Gauss.

    Local nodes: (0,0,0) (1,0,0) (1,1,0) (0,1,0) (0,0,1) (1,0,1) (1,1,1) (0,1,1);
    dofs per node (ux, uy, uz).  Voigt order xx, yy, zz, xy, yz, zx with
    engineering shear strains.
    """
    xi_n = np.array([-1, 1, 1, -1, -1, 1, 1, -1], dtype=float)
    et_n = np.array([-1, -1, 1, 1, -1, -1, 1, 1], dtype=float)
    ze_n = np.array([-1, -1, -1, -1, 1, 1, 1, 1], dtype=float)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] += 2 * mu
    D[3, 3] = D[4, 4] = D[5, 5] = mu
    K = np.zeros((24, 24))
    jac = h / 2.0  # dx/dxi for an axis-aligned cube
    for xi in _GAUSS2:
        for et in _GAUSS2:
            for ze in _GAUSS2:
                dN = np.empty((3, 8))
                dN[0] = xi_n * (1 + et * et_n) * (1 + ze * ze_n) / 8.0
                dN[1] = et_n * (1 + xi * xi_n) * (1 + ze * ze_n) / 8.0
                dN[2] = ze_n * (1 + xi * xi_n) * (1 + et * et_n) / 8.0
                dN /= jac
                B = np.zeros((6, 24))
                B[0, 0::3] = dN[0]
                B[1, 1::3] = dN[1]
                B[2, 2::3] = dN[2]
                B[3, 0::3] = dN[1]
                B[3, 1::3] = dN[0]
                B[4, 1::3] = dN[2]
                B[4, 2::3] = dN[1]
                B[5, 0::3] = dN[2]
                B[5, 2::3] = dN[0]
                K += B.T @ D @ B * jac ** 3
    return 0.5 * (K + K.T)


_HEX8_NODES = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                        [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)


def hex8_stiffness_at(X: np.ndarray, E: float = 1.0, nu: float = 0.3) -> np.ndarray:
    """24x24 stiffness of a trilinear hexahedron with node coordinates X (8, 3)
    in the local node order of hex8_stiffness: sum over the 2x2x2 Gauss points
    of det(J) B' D B, J = dN/dxi X, dN/dx = J^-1 dN/dxi (isoparametric map)."""
    X = np.asarray(X, dtype=float)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] += 2 * mu
    D[3, 3] = D[4, 4] = D[5, 5] = mu
    s = _HEX8_NODES
    K = np.zeros((24, 24))
    for xi in _GAUSS2:
        for et in _GAUSS2:
            for ze in _GAUSS2:
                dN = np.empty((3, 8))
                dN[0] = s[:, 0] * (1 + et * s[:, 1]) * (1 + ze * s[:, 2]) / 8.0
                dN[1] = s[:, 1] * (1 + xi * s[:, 0]) * (1 + ze * s[:, 2]) / 8.0
                dN[2] = s[:, 2] * (1 + xi * s[:, 0]) * (1 + et * s[:, 1]) / 8.0
                J = dN @ X
                dx = np.linalg.solve(J, dN)
                B = np.zeros((6, 24))
                B[0, 0::3] = dx[0]
                B[1, 1::3] = dx[1]
                B[2, 2::3] = dx[2]
                B[3, 0::3] = dx[1]
                B[3, 1::3] = dx[0]
                B[4, 1::3] = dx[2]
                B[4, 2::3] = dx[1]
                B[5, 0::3] = dx[2]
                B[5, 2::3] = dx[0]
                K += B.T @ D @ B * np.linalg.det(J)
    return K


def hex8_coordinates(n: int, distort: float = 0.0, seed: int = 0) -> np.ndarray:
    """(nanod, 3) node coordinates of hex8_grid(n): node i + (n+1)(j + (n+1)k)
    at (i, j, k), unit elements; with ``distort`` every node is moved by
    U[-distort, distort] per coordinate (default_rng(seed)), nodes on the
    clamped face z = 0 included."""
    m = n + 1
    k, j, i = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    xyz = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(float)
    if distort:
        xyz += np.random.default_rng(seed).uniform(-distort, distort, xyz.shape)
    return xyz


def quad4_stiffness(E: float = 1.0, nu: float = 0.25, t: float = 1.0, h: float = 1.0) -> np.ndarray:
    """8x8 plane-stress stiffness of a bilinear square (edge h, thickness t), 2x2 Gauss.

    Local nodes (0,0) (1,0) (1,1) (0,1); dofs (ux, uy); Voigt xx, yy, xy.
    """
    xi_n = np.array([-1, 1, 1, -1], dtype=float)
    et_n = np.array([-1, -1, 1, 1], dtype=float)
    c = E / (1 - nu * nu)
    D = c * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])
    K = np.zeros((8, 8))
    jac = h / 2.0
    for xi in _GAUSS2:
        for et in _GAUSS2:
            dN = np.empty((2, 4))
            dN[0] = xi_n * (1 + et * et_n) / 4.0
            dN[1] = et_n * (1 + xi * xi_n) / 4.0
            dN /= jac
            B = np.zeros((3, 8))
            B[0, 0::2] = dN[0]
            B[1, 1::2] = dN[1]
            B[2, 0::2] = dN[1]
            B[2, 1::2] = dN[0]
            K += B.T @ D @ B * t * jac ** 2
    return 0.5 * (K + K.T)


def beam3d_stiffness(p0: np.ndarray, p1: np.ndarray, EA: float, EI: float, GJ: float) -> np.ndarray:
    """12x12 global stiffness of a 3D Euler-Bernoulli beam between p0 and p1.

    EIy = EIz = EI.  Dofs per node: ux, uy, uz, rx, ry, rz.  The local frame has
    x along the member; y is perpendicular to x and to the global axis least
    aligned with x.
    """
    d = np.asarray(p1, float) - np.asarray(p0, float)
    L = float(np.linalg.norm(d))
    ex = d / L
    ref = np.eye(3)[int(np.argmin(np.abs(ex)))]
    ey = np.cross(ref, ex)
    ey /= np.linalg.norm(ey)
    ez = np.cross(ex, ey)
    R = np.vstack([ex, ey, ez])
    k = np.zeros((12, 12))
    a = EA / L
    k[0, 0] = k[6, 6] = a
    k[0, 6] = k[6, 0] = -a
    g = GJ / L
    k[3, 3] = k[9, 9] = g
    k[3, 9] = k[9, 3] = -g
    # bending about local z (deflection v = dof 1/7, rotation rz = dof 5/11)
    # and about local y (deflection w = dof 2/8, rotation ry = dof 4/10)
    for (v1, r1, v2, r2, s) in ((1, 5, 7, 11, 1.0), (2, 4, 8, 10, -1.0)):
        c1 = 12 * EI / L ** 3
        c2 = 6 * EI / L ** 2
        c3 = 4 * EI / L
        c4 = 2 * EI / L
        idx = [v1, r1, v2, r2]
        kb = np.array([[c1, s * c2, -c1, s * c2],
                       [s * c2, c3, -s * c2, c4],
                       [-c1, -s * c2, c1, -s * c2],
                       [s * c2, c4, -s * c2, c3]])
        k[np.ix_(idx, idx)] += kb
    T = np.zeros((12, 12))
    for b in range(4):
        T[3 * b:3 * b + 3, 3 * b:3 * b + 3] = R
    K = T.T @ k @ T
    return 0.5 * (K + K.T)


def beam3d_stiffness_batch(p0: np.ndarray, p1: np.ndarray, EA: np.ndarray, EI: np.ndarray,
                           GJ: np.ndarray) -> np.ndarray:
    """beam3d_stiffness for many members at once: p0, p1 (n, 3), EA, EI, GJ
    (n,); returns (n, 12, 12), member e equal to beam3d_stiffness(p0[e],
    p1[e], EA[e], EI[e], GJ[e]) up to rounding."""
    d = np.asarray(p1, float) - np.asarray(p0, float)
    n = d.shape[0]
    L = np.sqrt(np.einsum("ij,ij->i", d, d))
    ex = d / L[:, None]
    ref = np.eye(3)[np.argmin(np.abs(ex), axis=1)]
    ey = np.cross(ref, ex)
    ey /= np.sqrt(np.einsum("ij,ij->i", ey, ey))[:, None]
    ez = np.cross(ex, ey)
    R = np.stack([ex, ey, ez], axis=1)  # (n, 3, 3), rows ex, ey, ez
    k = np.zeros((n, 12, 12))
    a = EA / L
    g = GJ / L
    for (i, j, v) in ((0, 0, a), (6, 6, a), (0, 6, -a), (6, 0, -a),
                      (3, 3, g), (9, 9, g), (3, 9, -g), (9, 3, -g)):
        k[:, i, j] = v
    c1, c2, c3, c4 = 12 * EI / L ** 3, 6 * EI / L ** 2, 4 * EI / L, 2 * EI / L
    for (v1, r1, v2, r2, sg) in ((1, 5, 7, 11, 1.0), (2, 4, 8, 10, -1.0)):
        idx = (v1, r1, v2, r2)
        kb = np.stack([np.stack([c1, sg * c2, -c1, sg * c2], -1),
                       np.stack([sg * c2, c3, -sg * c2, c4], -1),
                       np.stack([-c1, -sg * c2, c1, -sg * c2], -1),
                       np.stack([sg * c2, c4, -sg * c2, c3], -1)], 1)
        for ii in range(4):
            for jj in range(4):
                k[:, idx[ii], idx[jj]] += kb[:, ii, jj]
    T = np.zeros((n, 12, 12))
    for b in range(4):
        T[:, 3 * b:3 * b + 3, 3 * b:3 * b + 3] = R
    K = np.matmul(np.matmul(T.transpose(0, 2, 1), k), T)
    return 0.5 * (K + K.transpose(0, 2, 1))


# --------------------------------------------------------------------------
# Meshes
# --------------------------------------------------------------------------

def hex8_grid(n: int, E: float = 1.0, nu: float = 0.3, e_sigma: float | None = None,
              e_seed: int | None = None) -> SamModel:
    """n x n x n unit-cube hex8 elements, face z=0 clamped.

    Nodes are numbered i + (n+1)*(j + (n+1)*k) + 1 (x fastest); elements
    likewise.  With ``e_sigma`` the element moduli are E * exp(N(0, e_sigma^2))
    drawn from ``default_rng(e_seed)`` in element order.
    """
    m = n + 1
    nanod = m ** 3
    nel = n ** 3
    ndof = 3 * nanod
    madof = (3 * np.arange(nanod + 1) + 1).astype(I8)
    k_of = np.repeat(np.arange(m), m * m)
    msc = np.ones(ndof, dtype=I8)
    fixed_nodes = np.nonzero(k_of == 0)[0]
    for c in range(3):
        msc[3 * fixed_nodes + c] = 0
    e = np.arange(nel)
    i, j, k = e % n, (e // n) % n, e // (n * n)  # x fastest
    base = i + m * (j + m * k)
    off = np.array([0, 1, 1 + m, m, m * m, 1 + m * m, 1 + m + m * m, m + m * m])
    conn = base[:, None] + off[None, :] + 1
    mmnpc = conn.ravel().astype(I8)
    mpmnpc = (8 * np.arange(nel + 1) + 1).astype(I8)
    K0 = hex8_stiffness(1.0, nu)
    escale = np.full(nel, E)
    meta = {"kind": "hex8", "n": n, "nu": nu, "E": E}
    if e_sigma is not None:
        rng = np.random.default_rng(e_seed)
        escale = E * np.exp(rng.normal(0.0, e_sigma, nel))
        meta.update(e_sigma=e_sigma, e_seed=e_seed)
    return SamModel(f"hex8_{n}", nanod, nel, ndof, madof, msc, mpmnpc, mmnpc,
                    np.full(nel, 24, dtype=I8), K0.ravel(order="F").copy(),
                    np.zeros(nel, dtype=I8), escale, meta)


def hex8_constrained(n: int, e_sigma: float | None = 0.5, e_seed: int | None = 1) -> SamModel:
    """hex8_grid(n) with constraint equations on the face x = 1 (nodes i = n,
    k >= 1), in node order and per node in dof order:

    * ux slave: 0.75 ux(n-1, j, k) + 0.25 ux(n-1, j', k) + 0.0 uz(n-1, j, k),
      j' = j+1 (j-1 on the edge j = n), c0 = 1e-3 k when j is a multiple of 3,
      else 0 (a master with coefficient zero, which the assembly skips);
    * uy slave (k < n): 0.5 uz(n, j, k), a master in the same node, c0 = 0;
    * uz prescribed (k = n): c0 = -0.01 for even j, else 0.

    The constrained dofs get status code 0 and the masters stay free, as SPRSEQ
    requires (sprseq.f:108-140).
    """
    mdl = hex8_grid(n, e_sigma=e_sigma, e_seed=e_seed)
    m = n + 1

    def dof(i, j, k, c):
        return 3 * (i + m * (j + m * k)) + c  # 0-based

    mp, mm, tt = [1], [], []

    def add(slave, c0, masters=()):
        mdl.msc[slave] = 0
        mm.append(slave + 1)
        tt.append(c0)
        for d, t in masters:
            mm.append(d + 1)
            tt.append(t)
        mp.append(len(mm) + 1)

    for k in range(1, m):
        for j in range(m):
            j2 = j + 1 if j < n else j - 1
            add(dof(n, j, k, 0), 1e-3 * k if j % 3 == 0 else 0.0,
                [(dof(n - 1, j, k, 0), 0.75), (dof(n - 1, j2, k, 0), 0.25),
                 (dof(n - 1, j, k, 2), 0.0)])
            if k < n:
                add(dof(n, j, k, 1), 0.0, [(dof(n, j, k, 2), 0.5)])
            else:
                add(dof(n, j, k, 2), -0.01 if j % 2 == 0 else 0.0)
    mdl.name = f"hex8c_{n}"
    mdl.meta.update(kind="hex8c")
    mdl.mpmceq = np.array(mp, dtype=I8)
    mdl.mmceq = np.array(mm, dtype=I8)
    mdl.ttcc = np.array(tt, dtype=np.float64)
    return mdl


def quad4_grid(n: int, E: float = 1.0, nu: float = 0.25, t: float = 1.0) -> SamModel:
    """n x n unit-square quad4 plane-stress elements, left edge (x=0) clamped."""
    m = n + 1
    nanod = m * m
    nel = n * n
    ndof = 2 * nanod
    madof = (2 * np.arange(nanod + 1) + 1).astype(I8)
    msc = np.ones(ndof, dtype=I8)
    left = np.arange(0, nanod, m)
    msc[2 * left] = 0
    msc[2 * left + 1] = 0
    j, i = np.divmod(np.arange(nel), n)
    base = i + m * j
    off = np.array([0, 1, 1 + m, m])
    conn = base[:, None] + off[None, :] + 1
    K0 = quad4_stiffness(1.0, nu, t)
    return SamModel(f"quad4_{n}", nanod, nel, ndof, madof, msc,
                    (4 * np.arange(nel + 1) + 1).astype(I8), conn.ravel().astype(I8),
                    np.full(nel, 8, dtype=I8), K0.ravel(order="F").copy(),
                    np.zeros(nel, dtype=I8), np.full(nel, E),
                    {"kind": "quad4", "n": n, "nu": nu, "E": E, "t": t})


def beam_frame(nnodes: int, rng: np.random.Generator, extra_ratio: float = 0.5,
               window: int = 8) -> SamModel:
    """A random 3D Euler-Bernoulli frame with ``nnodes`` nodes, 6 dofs each.

    Connectivity: a random spanning tree in which node k>0 joins a node chosen
    uniformly among the ``window`` nodes before it, plus
    ``round(extra_ratio*nnodes)`` extra members, each joining a random node a
    to a random node b with 0 < |a-b| <= window that is not yet joined to it.  Member lengths are
    U[0.5, 2]: each tree member leaves its parent in a random direction with
    that length; extra members take the distance between their end nodes.
    EA and EI are log-normal(0, 0.3) per member; GJ = EI / 1.3 (a solid
    circular section with nu = 0.3).  Node 1 is clamped.
    """
    pos = np.zeros((nnodes, 3))
    edges = []
    for kk in range(1, nnodes):
        par = int(rng.integers(max(0, kk - window), kk))
        dvec = rng.normal(size=3)
        dvec /= np.linalg.norm(dvec)
        pos[kk] = pos[par] + rng.uniform(0.5, 2.0) * dvec
        edges.append((par, kk))
    have = set(edges)
    n_extra = int(round(extra_ratio * nnodes))
    tries = 0
    while n_extra > 0 and tries < 50 * nnodes:
        tries += 1
        a = int(rng.integers(0, nnodes))
        b = a + int(rng.integers(-window, window + 1))
        if a == b or b < 0 or b >= nnodes:
            continue
        e = (min(a, b), max(a, b))
        if e in have:
            continue
        have.add(e)
        edges.append(e)
        n_extra -= 1
    nel = len(edges)
    EA = np.exp(rng.normal(0.0, 0.3, nel))
    EI = np.exp(rng.normal(0.0, 0.3, nel))
    ea = np.array(edges, dtype=np.int64).reshape(-1, 2)
    K = beam3d_stiffness_batch(pos[ea[:, 0]], pos[ea[:, 1]], EA, EI, EI / 1.3)
    em = np.ascontiguousarray(K.transpose(0, 2, 1)).ravel()  # each matrix column-major
    madof = (6 * np.arange(nnodes + 1) + 1).astype(I8)
    msc = np.ones(6 * nnodes, dtype=I8)
    msc[:6] = 0
    conn = (np.array(edges, dtype=I8) + 1).ravel()
    return SamModel(f"frame_{nnodes}", nnodes, nel, 6 * nnodes, madof, msc,
                    (2 * np.arange(nel + 1) + 1).astype(I8), conn,
                    np.full(nel, 12, dtype=I8), em, (144 * np.arange(nel)).astype(I8), None,
                    {"kind": "frame", "nnodes": nnodes, "pos": pos})


def uniform_loads(ndof: int, rng: np.random.Generator, nrhs: int = 1) -> np.ndarray:
    """Loads U[-1,1] per dof in natural dof order, shape (ndof, nrhs)."""
    return rng.uniform(-1.0, 1.0, (nrhs, ndof)).T.copy()


def normal_loads(ndof: int, rng: np.random.Generator, nrhs: int) -> np.ndarray:
    """Loads N(0,1) per dof in natural dof order, column by column, shape (ndof, nrhs)."""
    return rng.normal(0.0, 1.0, (nrhs, ndof)).T.copy()
