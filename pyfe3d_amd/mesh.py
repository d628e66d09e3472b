"""Structured plate meshes in the reference's conventions.

Every plate test and tutorial of pyfe3d builds its mesh the same way
(tests/test_quad4_static_point_load.py:25-40,53-78): ``linspace`` grids
with x-major node order (node position = i*ny + j, nx and ny counting
nodes), elements in ``nids_mesh[:-1, :-1].flatten()`` order with
n1=(i, j), n2=(i+1, j), n3=(i+1, j+1), n4=(i, j+1), DOF offsets
c = 6*position and slots ``init_k = SIZE*e``.  Triangles split each quad
into (n1, n2, n3) and (n1, n3, n4) (tests/test_tria3r_static_point_load.py).
"""
import numpy as np

DOF = 6


class PlateMesh:
    """Nodes, connectivity and boundary masks of a structured plate."""

    def __init__(self, nx, ny, a, b, jitter=0.0, seed=None, independent_xy=True):
        self.nx, self.ny, self.a, self.b = nx, ny, a, b
        xt = np.linspace(0, a, nx)
        yt = np.linspace(0, b, ny)
        xm, ym = np.meshgrid(xt, yt)
        x = xm.T.flatten()
        y = ym.T.flatten()
        if jitter:
            rng = np.random.default_rng(seed)
            inner = ~(np.isclose(x, 0) | np.isclose(x, a) | np.isclose(y, 0) | np.isclose(y, b))
            dx, dy = a / (nx - 1), b / (ny - 1)
            jx = rng.uniform(-jitter, jitter, inner.sum())
            jy = rng.uniform(-jitter, jitter, inner.sum()) if independent_xy else jx
            x = x.copy(); y = y.copy()
            x[inner] += dx * jx
            y[inner] += dy * jy
        self.ncoords = np.vstack((x, y, np.zeros_like(x))).T
        self.x = np.ascontiguousarray(self.ncoords.flatten())
        nid = np.arange(nx * ny).reshape(nx, ny)
        self.quads = np.stack([nid[:-1, :-1].flatten(), nid[1:, :-1].flatten(),
                               nid[1:, 1:].flatten(), nid[:-1, 1:].flatten()], axis=1)
        self.n_nodes = nx * ny
        self.n_dof = DOF * self.n_nodes

    @property
    def trias(self):
        q = self.quads
        t = np.empty((2 * q.shape[0], 3), dtype=q.dtype)
        t[0::2] = q[:, [0, 1, 2]]
        t[1::2] = q[:, [0, 2, 3]]
        return t

    def edge_nodes(self):
        x, y = self.ncoords[:, 0], self.ncoords[:, 1]
        return (np.isclose(x, 0.) | np.isclose(x, self.a) | np.isclose(y, 0.) |
                np.isclose(y, self.b))

    def clamped(self):
        """All 6 DOFs fixed at the boundary nodes."""
        fixed = np.zeros(self.n_dof, dtype=bool)
        e = self.edge_nodes()
        for d in range(DOF):
            fixed[d::DOF] = e
        return fixed

    def centre_node(self):
        x, y = self.ncoords[:, 0], self.ncoords[:, 1]
        return int(np.argmin((x - self.a / 2) ** 2 + (y - self.b / 2) ** 2))

    def pressure_load(self, p):
        """Consistent nodal forces f_z = sum_e int p N_I dA of a uniform
        pressure (bilinear shape functions, 2x2 Gauss), so sum f_z = p*area."""
        f = np.zeros(self.n_dof)
        q = self.quads
        xy = self.ncoords[:, :2]
        g = 1 / np.sqrt(3.)
        for xi in (-g, g):
            for eta in (-g, g):
                N = 0.25 * np.array([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta),
                                     (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)])
                dxi = 0.25 * np.array([-(1 - eta), (1 - eta), (1 + eta), -(1 + eta)])
                deta = 0.25 * np.array([-(1 - xi), -(1 + xi), (1 + xi), (1 - xi)])
                X = xy[q]  # (E, 4, 2)
                J11 = X[:, :, 0] @ dxi; J12 = X[:, :, 1] @ dxi
                J21 = X[:, :, 0] @ deta; J22 = X[:, :, 1] @ deta
                det = J11 * J22 - J12 * J21
                np.add.at(f, DOF * q + 2, p * det[:, None] * N[None, :])
        return f
