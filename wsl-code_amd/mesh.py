"""Meshes: the reference's StraightMeshReader interface plus the seeded
Voronoi generator (disk / annulus with circular boundary arcs)."""
import ctypes

import numpy as np

from ._lib import MeshOpts, VemdError, c_dp, c_ip, c_vp, check, lib, DOMAIN_DISK


def _d(a):
    return a.ctypes.data_as(c_dp)


def _i(a):
    return a.ctypes.data_as(c_ip)


class Mesh:
    """Owns a vemd_mesh handle.  Arrays follow vem::StraightMeshData
    (CPU/mesh/straight_mesh.h:18-77)."""

    def __init__(self, handle):
        self.h = c_vp(handle) if not isinstance(handle, c_vp) else handle
        c = np.zeros(5, np.int64)
        check(lib().vemd_mesh_counts(self.h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        self.num_nodes, self.num_cells, self.num_edges, self._ntot, self.num_arcs = (int(v) for v in c)

    @classmethod
    def concat(cls, meshes):
        """Disjoint union of meshes: a batch of independent instances."""
        arr = (c_vp * len(meshes))(*[m.h.value for m in meshes])
        h = c_vp()
        check(lib().vemd_mesh_concat(len(meshes), arr, ctypes.byref(h)))
        out = cls(h)
        out.instance_cells = np.cumsum([0] + [m.num_cells for m in meshes])
        out.parts = list(meshes)
        return out

    @classmethod
    def read(cls, path, vtk=False):
        h = c_vp()
        check(lib().vemd_mesh_read(str(path).encode(), 1 if vtk else 0, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, xy, cell_off, cell_nodes, arc_nodes=None, arc_radius=None):
        xy = np.ascontiguousarray(xy, np.float64).reshape(-1)
        off = np.ascontiguousarray(cell_off, np.int32)
        nodes = np.ascontiguousarray(cell_nodes, np.int32)
        if arc_nodes is None:
            arc_nodes, arc_radius = np.zeros(0, np.int32), np.zeros(0)
        an = np.ascontiguousarray(arc_nodes, np.int32).reshape(-1)
        ar = np.ascontiguousarray(arc_radius, np.float64).reshape(-1)
        # the C ABI reads cell_nodes up to cell_off[-1] and two nodes per arc
        if len(xy) % 2 or off.ndim != 1 or len(off) < 2 or off[0] != 0 or np.any(np.diff(off) < 0) \
                or off[-1] != len(nodes) or len(an) != 2 * len(ar):
            raise VemdError(1, "inconsistent mesh arrays: %d coordinates, %d offsets ending at %s, %d cell nodes, "
                               "%d arc nodes for %d radii" % (len(xy), len(off), off[-1] if len(off) else None,
                                                               len(nodes), len(an), len(ar)))
        h = c_vp()
        check(lib().vemd_mesh_from_arrays(len(xy) // 2, _d(xy), len(off) - 1, _i(off), _i(nodes), len(ar), _i(an),
                                          _d(ar), ctypes.byref(h)))
        return cls(h)

    def arrays(self):
        """Exports the mesh as a dict of numpy arrays."""
        nn, nc, ne, nt, na = self.num_nodes, self.num_cells, self.num_edges, self._ntot, self.num_arcs
        out = dict(xy=np.zeros(2 * nn), cell_off=np.zeros(nc + 1, np.int32), cell_nodes=np.zeros(nt, np.int32),
                   cell_edges=np.zeros(nt, np.int32), edge_nodes=np.zeros(2 * ne, np.int32),
                   edge_occ=np.zeros(ne, np.int32), cx=np.zeros(nc), cy=np.zeros(nc), area=np.zeros(nc),
                   diam=np.zeros(nc), arc_nodes=np.zeros(2 * max(na, 1), np.int32), arc_radius=np.zeros(max(na, 1)))
        check(lib().vemd_mesh_export(self.h, _d(out["xy"]), _i(out["cell_off"]), _i(out["cell_nodes"]),
                                     _i(out["cell_edges"]), _i(out["edge_nodes"]), _i(out["edge_occ"]), _d(out["cx"]),
                                     _d(out["cy"]), _d(out["area"]), _d(out["diam"]), _i(out["arc_nodes"]),
                                     _d(out["arc_radius"])))
        out["xy"] = out["xy"].reshape(-1, 2)
        out["arc_nodes"] = out["arc_nodes"][: 2 * na].reshape(-1, 2)
        out["arc_radius"] = out["arc_radius"][:na]
        return out

    def write_vtk(self, path):
        check(lib().vemd_mesh_write_vtk(self.h, str(path).encode()))

    def __del__(self):
        try:
            lib().vemd_mesh_destroy(self.h)
        except Exception:
            pass


def generate_mesh(num_cells, domain=DOMAIN_DISK, r_outer=1.0, r_inner=0.5, lloyd_iters=20, seed=1,
                  min_edge_frac=0.0):
    """Seeded Lloyd-relaxed Voronoi mesh of a disk or annulus with arc boundary edges."""
    o = MeshOpts(domain, r_inner, r_outer, num_cells, lloyd_iters, seed, min_edge_frac)
    h = c_vp()
    check(lib().vemd_mesh_generate(ctypes.byref(o), ctypes.byref(h)))
    return Mesh(h)


class StraightMeshReader:
    """Mirror of vem::StraightMeshReader (CPU/mesh/straight_mesh.h:80-143)."""

    def __init__(self):
        self.mesh = None

    def read_mesh(self, filename):
        self.mesh = Mesh.read(filename, vtk=False)
        return True

    def read_mesh_vtk(self, filename):
        self.mesh = Mesh.read(filename, vtk=True)
        return True

    def get_mesh_data(self):
        return self.mesh.arrays()
