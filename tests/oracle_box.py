"""A single-level box domain for the oracle, built here from first principles
(every block present; neighbour table and block pointer from the block
coordinates) so that pins of the oracle's pull, boundary cascade and thermal
step do not depend on the package's own domain builder."""
import ctypes as C

import numpy as np

from oracle import OrcParams, lib

K = np.arange(27)
CX, CY, CZ = K % 3 - 1, (K // 3) % 3 - 1, K // 9 - 1
MIRROR_Y = CX + 1 + 3 * (1 - CY) + 9 * (CZ + 1)   # cy -> -cy
MIRROR_Z = CX + 1 + 3 * (CY + 1) + 9 * (1 - CZ)   # cz -> -cz
OPP = 26 - K


def params(**kw):
    p = OrcParams()
    p.c_wale, p.nu_sgs_bg, p.inlet_turbulence = 0.0, 0.0, 0.0
    p.symmetric = p.wall_model = p.sponge_blend = 0
    p.temporal = 1
    p.omega_bulk = p.omega_3 = p.omega_4 = 1.0
    p.lambda_param = 1.0 / 6.0
    p.adaptive_omega4 = 0
    p.limiter = 0
    p.compressible = 0
    p.q_min = 0.001
    p.ddf = 0
    p.T_inlet = p.T_wall = 1.0
    p.wall_bc_adiabatic = 1
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Box:
    """nbx x nby x nbz blocks of 8^3 cells, one level, FP64 oracle by default.
    Fields are exchanged in the reference layout f[x,y,z,b,k] (column-major,
    1-based), here as arrays indexed [k, gx-1, gy-1, gz-1] for convenience."""

    def __init__(self, nbx, nby, nbz, tau=0.6, tau_g=0.6, thermal=False, precision="f64", obstacle=None):
        self.L = lib(precision)
        self.shape = (8 * nbx, 8 * nby, 8 * nbz)
        coords = np.array([(bx, by, bz) for bz in range(1, nbz + 1) for by in range(1, nby + 1)
                           for bx in range(1, nbx + 1)], dtype=np.int32)
        nb = len(coords)
        self.nb = nb
        self.coords = coords
        ptr = np.zeros((nbx, nby, nbz), dtype=np.int32)
        for i, (bx, by, bz) in enumerate(coords):
            ptr[bx - 1, by - 1, bz - 1] = i + 1
        nbt = np.zeros((27, nb), dtype=np.int32)           # (nb,27) column-major
        for i, (bx, by, bz) in enumerate(coords):
            for d in range(27):
                x, y, z = bx + d % 3 - 1, by + (d // 3) % 3 - 1, bz + d // 9 - 1
                if 1 <= x <= nbx and 1 <= y <= nby and 1 <= z <= nbz:
                    nbt[d, i] = ptr[x - 1, y - 1, z - 1]
        n4 = nb * 512
        # cell index of global (gx, gy, gz), 0-based arrays
        gx, gy, gz = np.meshgrid(np.arange(self.shape[0]), np.arange(self.shape[1]), np.arange(self.shape[2]),
                                 indexing="ij")
        b = ptr[gx // 8, gy // 8, gz // 8] - 1
        self.lin = (gx % 8) + 8 * (gy % 8) + 64 * (gz % 8) + 512 * b   # linear cell of each global cell
        obst = np.zeros(n4, dtype=np.uint8)
        if obstacle is not None:
            obst[self.lin[obstacle]] = 1
        self.obst = obst
        self._keep = [coords, nbt, ptr, obst]
        self.ctx = self.L.orc_create(1, *self.shape)
        mx, my, mz = (np.ascontiguousarray(coords[:, i]) for i in range(3))
        bp = np.ascontiguousarray(ptr.ravel(order="F"))
        sp = np.zeros(n4, dtype=np.float32)
        wd = np.full(n4, 100.0, dtype=np.float32)
        self._keep += [mx, my, mz, bp, sp, wd]
        ip = C.c_int
        rc = self.L.orc_set_level(self.ctx, 0, nb, _p(mx, ip), _p(my, ip), _p(mz, ip), _p(nbt, ip), _p(bp, ip),
                                  nbx, nby, nbz, _p(obst, C.c_uint8), _p(sp, C.c_float), _p(wd, C.c_float),
                                  float(tau), float(tau_g), 1, int(thermal), 0, None, None, None, None, None)
        assert rc == 0
        self.t = 0

    def close(self):
        if self.ctx:
            self.L.orc_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        self.close()

    def init(self, T0=1.0):
        self.L.orc_init_equilibrium(self.ctx, float(T0))
        self.t = 0

    def step(self, prm, u_in, n=1):
        self.L.orc_step_batch(self.ctx, self.t + 1, n, float(u_in), C.byref(prm))
        self.t += n

    def _field(self, which, ncomp):
        out = np.zeros(self.nb * 512 * ncomp)
        assert self.L.orc_get(self.ctx, 0, which, self.t, _p(out, C.c_double)) == 0
        a = out.reshape(ncomp, self.nb * 512)
        return a[:, self.lin]

    def get(self, field):
        which, n = {"f": (0, 27), "rho": (1, 1), "vel": (2, 3), "g": (3, 27), "T": (4, 1)}[field]
        return self._field(which, n)

    def put(self, field, grid):
        """grid: [ncomp, nx, ny, nz]; written at both parities."""
        which, n = {"f": (0, 27), "rho": (1, 1), "vel": (2, 3), "g": (3, 27), "T": (4, 1)}[field]
        flat = np.zeros((n, self.nb * 512))
        flat[:, self.lin] = np.asarray(grid, dtype=np.float64).reshape(n, *self.shape)
        flat = np.ascontiguousarray(flat.ravel())
        for t in (0, 1):
            assert self.L.orc_put(self.ctx, 0, which, t, _p(flat, C.c_double)) == 0


def feq(L, rho, u, k, comp=0):
    return L.orc_feq(comp, rho, u[0], u[1], u[2], int(k))


def weights():
    d2 = CX * CX + CY * CY + CZ * CZ
    return np.where(d2 == 0, 8 / 27, np.where(d2 == 1, 2 / 27, np.where(d2 == 2, 1 / 54, 1 / 216)))
