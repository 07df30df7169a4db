"""Exodus II mesh and solution files for the Newton path.

Goma takes its mesh from an Exodus II file (rd_mesh.c:152 -> rd_exo.c) and
writes results into one (wr_exo.c:89 for the mesh, wr_soln.c:40-131 for the
fields).  Exodus II is a naming convention over netCDF; this module writes
and reads it through ``goma_amd.netcdf3`` (64-bit-offset netCDF by default,
the classic netCDF family; netCDF-4/HDF5 Exodus files are not read).

What is kept from the Exodus II layout Goma sees:
  * dimensions ``num_dim``, ``num_nodes``, ``num_elem``, ``num_el_blk``,
    ``num_el_in_blk<b>``, ``num_nod_per_el<b>``, ``num_node_sets``,
    ``num_nod_ns<k>``, ``num_side_sets``, ``num_side_ss<k>``, ``time_step`` (unlimited), ``len_name``, ``len_string``,
    ``len_line``, ``four``;
  * coordinates ``coordx/coordy[/coordz]`` (or the older ``coord[num_dim]``),
    block connectivity ``connect<b>`` (1-based, attribute ``elem_type``
    QUAD9 / HEX27), block and node-set IDs in ``eb_prop1`` / ``ns_prop1``
    (attribute ``name = "ID"``), node-set members ``node_ns<k>`` (1-based),
    side sets ``ss_prop1``, ``elem_ss<k>`` (1-based) and ``side_ss<k>``
    (Exodus side numbers, ``SIDE_CORNERS``);
  * results: ``time_whole``, nodal variables ``name_nod_var`` /
    ``vals_nod_var<i>`` and element variables ``name_elem_var`` /
    ``vals_elem_var<i>eb<b>`` with ``elem_var_tab``.

Field names are Goma's (``Exo_Var_Names``, include/mm_names.h:7943-7952):
nodal VX, VY, VZ, T; the discontinuous P1 pressure is written the way Goma
writes an I_P1 variable, as element variables P_E1 .. P_E<dim+1>, one per
coefficient of {1, s, t[, u]} (src/mm_post_proc.c:12346-12358,
src/wr_soln.c:110-131).  QUAD9 and HEX27 local node order in Exodus is the
order Goma's shape functions use (src/rf_shape.c), so connectivity is taken
as it is.
"""
from __future__ import annotations

import numpy as np

from . import netcdf3 as nc
from .mesh import CENTROID_LOCAL, Mesh, band_order
from .mesh import renumber as renumber_mesh
from .problems import PRES, TEMP, Problem, _BC

ELEM_TYPE = {2: "QUAD9", 3: "HEX27"}
NODES_PER_ELEM = {2: 9, 3: 27}
NODAL_NAMES = ("VX", "VY", "VZ")
LEN_NAME = 33  # maximum_name_length 32 + NUL

# Exodus II side numbering: local nodes (0-based, corner nodes only) of each
# side.  QUAD: side 1 = nodes 1-2 (t = -1), 2 = 2-3 (s = +1), 3 = 3-4
# (t = +1), 4 = 4-1 (s = -1).  HEX: side 1 = 1,2,6,5 (t = -1), 2 = 2,3,7,6
# (s = +1), 3 = 3,4,8,7 (t = +1), 4 = 1,5,8,4 (s = -1), 5 = 1,4,3,2
# (u = -1), 6 = 5,6,7,8 (u = +1).
SIDE_CORNERS = {
    2: ((0, 1), (1, 2), (2, 3), (3, 0)),
    3: ((0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (0, 4, 7, 3), (0, 3, 2, 1), (4, 5, 6, 7)),
}


def wall_node_sets(mesh: Mesh) -> dict:
    """Node sets of the faces of a structured unit square / cube:
    ID 1, 2 = x = 0, 1; ID 3, 4 = y = 0, 1; ID 5, 6 = z = 0, 1."""
    lat = mesh.lattice()
    m = 2 * mesh.n
    out = {}
    for k in range(mesh.dim):
        out[2 * k + 1] = np.flatnonzero(lat[:, k] == 0)
        out[2 * k + 2] = np.flatnonzero(lat[:, k] == m)
    return out


def side_nodes(mesh: Mesh, elems, sides) -> np.ndarray:
    """Nodes of element sides (``sides`` 1-based, Exodus numbering): the
    side's corner nodes plus every Q2 node whose reference position lies on
    the same face, as an array [len(elems), 3 (QUAD9) or 9 (HEX27)]."""
    from .mesh import HEX27_POS, QUAD9_POS

    dim = mesh.dim
    pos = QUAD9_POS if dim == 2 else HEX27_POS
    tables = []
    for corners in SIDE_CORNERS[dim]:
        c = pos[list(corners)]
        axis = int(np.flatnonzero(np.all(c == c[0], axis=0))[0])  # the coordinate fixed on this side
        tables.append(np.flatnonzero(pos[:, axis] == c[0, axis]))
    tab = np.array(tables)
    sides = np.asarray(sides, np.int64)
    if sides.size and (sides.min() < 1 or sides.max() > len(tables)):
        raise ValueError(f"side numbers run 1..{len(tables)}")
    return mesh.connect[np.asarray(elems, np.int64)[:, None], tab[sides - 1]]


def wall_side_sets(mesh: Mesh) -> dict:
    """Side sets of the faces of a structured unit square / cube, with the
    IDs of ``wall_node_sets``: ID -> (elements 0-based, sides 1-based)."""
    lat = mesh.lattice()
    m = 2 * mesh.n
    out = {}
    dim = mesh.dim
    for sid in range(1, len(SIDE_CORNERS[dim]) + 1):
        nodes = side_nodes(mesh, np.arange(mesh.num_elems), np.full(mesh.num_elems, sid))
        for k in range(dim):
            for face, val in ((2 * k + 1, 0), (2 * k + 2, m)):
                on = np.all(lat[nodes, k] == val, axis=1)
                if on.any():
                    e = np.flatnonzero(on)
                    prev = out.get(face, (np.zeros(0, np.int64), np.zeros(0, np.int64)))
                    out[face] = (np.concatenate([prev[0], e]), np.concatenate([prev[1], np.full(e.size, sid)]))
    for face, (e, sd) in out.items():
        order = np.lexsort((sd, e))
        out[face] = (e[order], sd[order])
    return out


def first_unknown(mesh: Mesh, energy: bool) -> np.ndarray:
    """Global row of each node's first unknown (Goma's node-major numbering:
    U1..Ud[, T] at every node, then the element's dim+1 P1 coefficients at
    its centroid node; src/mm_fill_util.c:3516).  Length num_nodes + 1."""
    dim = mesh.dim
    count = np.full(mesh.num_nodes, dim + int(bool(energy)), np.int64)
    count[mesh.connect[:, CENTROID_LOCAL[dim]]] += dim + 1
    out = np.zeros(mesh.num_nodes + 1, np.int64)
    np.cumsum(count, out=out[1:])
    return out


def solution_fields(prob: Problem, x: np.ndarray):
    """Split a solution vector into Goma's output fields.

    Returns (nodal, elemental): dicts name -> array; nodal arrays have one
    value per node (VX, VY[, VZ][, T]), elemental ones one value per element
    (P_E1 .. P_E<dim+1>)."""
    mesh, dim, en = prob.mesh, prob.mesh.dim, int(prob.phys.energy)
    fu = first_unknown(mesh, en)
    x = np.asarray(x, np.float64)
    if x.shape != (fu[-1],):
        raise ValueError(f"solution has {x.shape} entries, the problem has {fu[-1]} unknowns")
    nodal = {NODAL_NAMES[a]: x[fu[:-1] + a] for a in range(dim)}
    if en:
        nodal["T"] = x[fu[:-1] + dim]
    p0 = fu[mesh.connect[:, CENTROID_LOCAL[dim]]] + dim + en
    elem = {f"P_E{k + 1}": x[p0 + k] for k in range(dim + 1)}
    return nodal, elem


def write(path, mesh: Mesh, node_sets: dict | None = None, times=None, nodal: dict | None = None,
          elem: dict | None = None, title: str = "goma_amd", version: int = 2,
          side_sets: dict | None = None) -> None:
    """Write a one-block Exodus II file: mesh, node sets, side sets (ID ->
    (elements 0-based, sides 1-based)) and (optionally) results at ``times``
    (``nodal[name]`` shaped [len(times), num_nodes], ``elem[name]`` shaped
    [len(times), num_elem])."""
    dim, ne, nn = mesh.dim, mesh.num_elems, mesh.num_nodes
    if mesh.nodes_per_elem != NODES_PER_ELEM[dim]:
        raise ValueError(f"{dim}D meshes here are {ELEM_TYPE[dim]}")
    node_sets = dict(node_sets or {})
    nodal, elem = dict(nodal or {}), dict(elem or {})
    times = np.atleast_1d(np.asarray([] if times is None else times, np.float64))
    if (nodal or elem) and times.size == 0:
        raise ValueError("results need at least one time value")
    dims = {"len_string": 33, "len_line": 81, "four": 4, "len_name": LEN_NAME, "time_step": None,
            "num_dim": dim, "num_nodes": nn, "num_elem": ne, "num_el_blk": 1,
            "num_el_in_blk1": ne, "num_nod_per_el1": mesh.nodes_per_elem}
    V = nc.Var
    vs = [V("time_whole", ("time_step",), times),
          V("eb_status", ("num_el_blk",), np.ones(1, np.int32)),
          V("eb_prop1", ("num_el_blk",), np.ones(1, np.int32), {"name": "ID"}),
          V("eb_names", ("num_el_blk", "len_name"), nc.chars([""], LEN_NAME))]
    if node_sets:
        ids = sorted(node_sets)
        dims["num_node_sets"] = len(ids)
        vs.append(V("ns_status", ("num_node_sets",), np.ones(len(ids), np.int32)))
        vs.append(V("ns_prop1", ("num_node_sets",), np.asarray(ids, np.int32), {"name": "ID"}))
        vs.append(V("ns_names", ("num_node_sets", "len_name"), nc.chars([""] * len(ids), LEN_NAME)))
    for k, c in enumerate("xyz"[:dim]):
        vs.append(V("coord" + c, ("num_nodes",), np.ascontiguousarray(mesh.coor[k], np.float64)))
    vs.append(V("coor_names", ("num_dim", "len_name"), nc.chars(list("xyz"[:dim]), LEN_NAME)))
    conn = mesh.connect.astype(np.int64)
    if conn.size and (conn.min() < 0 or conn.max() >= nn):
        raise ValueError("connectivity refers to nodes outside the mesh")
    vs.append(V("connect1", ("num_el_in_blk1", "num_nod_per_el1"), (conn + 1).astype(np.int32),
                {"elem_type": ELEM_TYPE[dim]}))
    if node_sets:
        for k, i in enumerate(sorted(node_sets), start=1):
            nodes = np.asarray(node_sets[i], np.int64)
            if nodes.size and (nodes.min() < 0 or nodes.max() >= nn):
                raise ValueError(f"node set {i} refers to nodes outside the mesh")
            if nodes.size == 0:
                raise ValueError(f"node set {i} is empty")
            dims[f"num_nod_ns{k}"] = nodes.size
            vs.append(V(f"node_ns{k}", (f"num_nod_ns{k}",), (nodes + 1).astype(np.int32)))
    if side_sets:
        ids = sorted(side_sets)
        dims["num_side_sets"] = len(ids)
        vs.append(V("ss_status", ("num_side_sets",), np.ones(len(ids), np.int32)))
        vs.append(V("ss_prop1", ("num_side_sets",), np.asarray(ids, np.int32), {"name": "ID"}))
        nsides = len(SIDE_CORNERS[dim])
        for k, i in enumerate(ids, start=1):
            e, sd = (np.asarray(a, np.int64) for a in side_sets[i])
            if e.size == 0 or e.shape != sd.shape:
                raise ValueError(f"side set {i}: needs matching, non-empty element and side lists")
            if e.min() < 0 or e.max() >= ne or sd.min() < 1 or sd.max() > nsides:
                raise ValueError(f"side set {i} refers to elements or sides outside the mesh")
            dims[f"num_side_ss{k}"] = e.size
            vs.append(V(f"elem_ss{k}", (f"num_side_ss{k}",), (e + 1).astype(np.int32)))
            vs.append(V(f"side_ss{k}", (f"num_side_ss{k}",), sd.astype(np.int32)))
    if nodal:
        dims["num_nod_var"] = len(nodal)
        vs.append(V("name_nod_var", ("num_nod_var", "len_name"), nc.chars(list(nodal), LEN_NAME)))
        for i, (name, a) in enumerate(nodal.items(), start=1):
            a = np.asarray(a, np.float64).reshape(times.size, nn)
            vs.append(V(f"vals_nod_var{i}", ("time_step", "num_nodes"), a))
    if elem:
        dims["num_elem_var"] = len(elem)
        vs.append(V("name_elem_var", ("num_elem_var", "len_name"), nc.chars(list(elem), LEN_NAME)))
        vs.append(V("elem_var_tab", ("num_el_blk", "num_elem_var"), np.ones((1, len(elem)), np.int32)))
        for i, (name, a) in enumerate(elem.items(), start=1):
            a = np.asarray(a, np.float64).reshape(times.size, ne)
            vs.append(V(f"vals_elem_var{i}eb1", ("time_step", "num_el_in_blk1"), a))
    attrs = {"api_version": np.float32(8.0), "version": np.float32(8.0),
             "floating_point_word_size": 8, "file_size": 1 if version == 2 else 0,
             "maximum_name_length": LEN_NAME - 1, "int64_status": 0, "title": title}
    nc.write(path, dims, vs, attrs, version=version)


def write_solution(path, prob: Problem, x: np.ndarray, time: float = 0.0,
                   node_sets: dict | None = None, title: str | None = None) -> None:
    """Mesh plus one time step of Goma's output fields for solution ``x``.

    A problem that ``problem_from_exodus(..., renumber=True)`` built is
    written in the numbering of the file it came from, with its node sets."""
    nodal, elem = solution_fields(prob, x)
    mesh = prob.mesh
    if "exodus_node_order" in prob.meta:
        no, eo = prob.meta["exodus_node_order"], prob.meta["exodus_elem_order"]
        for k, v in nodal.items():
            nodal[k] = np.empty_like(v)
            nodal[k][no] = v
        for k, v in elem.items():
            elem[k] = np.empty_like(v)
            elem[k][eo] = v
        mesh = prob.meta["exodus_file_mesh"]
        if node_sets is None:
            node_sets = prob.meta.get("exodus_node_sets")
    write(path, mesh, node_sets=node_sets, times=[time],
          nodal={k: v[None] for k, v in nodal.items()}, elem={k: v[None] for k, v in elem.items()},
          title=title or prob.name)


def append_solution(path, prob: Problem, x: np.ndarray, time: float) -> int:
    """One more time step of Goma's output fields for solution ``x``,
    appended to a file ``write_solution`` made for the same problem (Goma
    writes each output step into the same results file, wr_soln.c:40-131).
    Returns the number of time steps now in the file."""
    nodal, elem = solution_fields(prob, x)
    if "exodus_node_order" in prob.meta:
        no, eo = prob.meta["exodus_node_order"], prob.meta["exodus_elem_order"]
        for k, v in nodal.items():
            nodal[k] = np.empty_like(v)
            nodal[k][no] = v
        for k, v in elem.items():
            elem[k] = np.empty_like(v)
            elem[k][eo] = v
    ds = nc.read(path, only={"name_nod_var", "name_elem_var", "time_whole"})
    v = ds.variables
    times = v["time_whole"].data if "time_whole" in v else np.zeros(0)
    if times.size and time <= times[-1]:
        raise ValueError(f"time {time} does not follow the file's last time {times[-1]}")
    rec = {"time_whole": np.array(float(time))}
    names = nc.strings(v["name_nod_var"].data) if "name_nod_var" in v else []
    if sorted(names) != sorted(nodal):
        raise ValueError(f"the file's nodal variables {names} are not this problem's {list(nodal)}")
    for i, name in enumerate(names, start=1):
        rec[f"vals_nod_var{i}"] = nodal[name]
    names = nc.strings(v["name_elem_var"].data) if "name_elem_var" in v else []
    if sorted(names) != sorted(elem):
        raise ValueError(f"the file's element variables {names} are not this problem's {list(elem)}")
    for i, name in enumerate(names, start=1):
        rec[f"vals_elem_var{i}eb1"] = elem[name]
    return nc.append_record(path, rec)


class ExodusFile:
    """Contents of an Exodus II file: ``mesh`` (0-based connectivity in
    Goma's local node order), ``node_sets`` (ID -> 0-based nodes),
    ``side_sets`` (ID -> (0-based elements, 1-based Exodus sides)), ``times``,
    ``nodal`` (name -> [time_step, num_nodes]) and ``elem`` (name ->
    [time_step, num_elem], blocks concatenated in block order)."""

    def __init__(self, path):
        ds = nc.read(path)
        v, d = ds.variables, ds.dims
        self.title = ds.attrs.get("title", "")
        dim = int(d["num_dim"])
        if dim not in ELEM_TYPE:
            raise ValueError(f"num_dim {dim}: only 2D QUAD9 and 3D HEX27 meshes are supported")
        nn = int(d["num_nodes"])
        if "coordx" in v:
            coor = np.stack([v["coord" + c].data for c in "xyz"[:dim]]).astype(np.float64)
        elif "coord" in v:
            coor = np.asarray(v["coord"].data, np.float64).reshape(dim, nn)
        else:
            raise ValueError("no coordinates in the file")
        nblk = int(d.get("num_el_blk", 0))
        conns, self.block_ids, self.block_sizes = [], [], []
        ids = v["eb_prop1"].data if "eb_prop1" in v else np.arange(1, nblk + 1)
        for b in range(1, nblk + 1):
            if f"connect{b}" not in v:  # a block with no elements has no connectivity
                self.block_ids.append(int(ids[b - 1]))
                self.block_sizes.append(0)
                continue
            cv = v[f"connect{b}"]
            et = str(cv.attrs.get("elem_type", "")).upper()
            npe = cv.data.shape[1]
            want = ELEM_TYPE[dim]
            if npe != NODES_PER_ELEM[dim] or not et.startswith(want[:3]):
                raise ValueError(f"block {ids[b - 1]}: element type {et} with {npe} nodes; "
                                 f"this path takes {want}")
            conns.append(np.asarray(cv.data, np.int64) - 1)
            self.block_ids.append(int(ids[b - 1]))
            self.block_sizes.append(cv.data.shape[0])
        if not conns:
            raise ValueError("the file has no elements")
        connect = np.concatenate(conns)
        if connect.min() < 0 or connect.max() >= nn:
            raise ValueError("connectivity refers to nodes outside the mesh")
        self.mesh = Mesh(dim=dim, n=0, coor=coor, connect=connect.astype(np.int32))
        self.node_sets = {}
        nns = int(d.get("num_node_sets", 0))
        nsid = v["ns_prop1"].data if "ns_prop1" in v else np.arange(1, nns + 1)
        for k in range(1, nns + 1):
            if f"node_ns{k}" in v:
                nodes = np.asarray(v[f"node_ns{k}"].data, np.int64) - 1
                if nodes.size and (nodes.min() < 0 or nodes.max() >= nn):
                    raise ValueError(f"node set {nsid[k - 1]} refers to nodes outside the mesh")
                self.node_sets[int(nsid[k - 1])] = nodes
        self.side_sets = {}
        nss = int(d.get("num_side_sets", 0))
        ssid = v["ss_prop1"].data if "ss_prop1" in v else np.arange(1, nss + 1)
        for k in range(1, nss + 1):
            if f"elem_ss{k}" in v:
                e = np.asarray(v[f"elem_ss{k}"].data, np.int64) - 1
                sd = np.asarray(v[f"side_ss{k}"].data, np.int64)
                if e.size and (e.min() < 0 or e.max() >= connect.shape[0]):
                    raise ValueError(f"side set {ssid[k - 1]} refers to elements outside the mesh")
                self.side_sets[int(ssid[k - 1])] = (e, sd)
        self.times = np.asarray(v["time_whole"].data, np.float64) if "time_whole" in v else np.zeros(0)
        nt = self.times.size
        self.nodal = {}
        if "name_nod_var" in v:
            names = nc.strings(v["name_nod_var"].data)
            for i, name in enumerate(names, start=1):
                if f"vals_nod_var{i}" in v:
                    self.nodal[name] = np.asarray(v[f"vals_nod_var{i}"].data, np.float64)
                elif "vals_nod_var" in v:  # older layout [time_step, num_nod_var, num_nodes]
                    self.nodal[name] = np.asarray(v["vals_nod_var"].data[:, i - 1], np.float64)
        self.elem = {}
        if "name_elem_var" in v:
            names = nc.strings(v["name_elem_var"].data)
            tab = (np.asarray(v["elem_var_tab"].data) if "elem_var_tab" in v
                   else np.ones((nblk, len(names)), np.int32))
            for i, name in enumerate(names, start=1):
                parts = []
                for b in range(1, nblk + 1):
                    key = f"vals_elem_var{i}eb{b}"
                    if tab[b - 1, i - 1] and key in v:
                        parts.append(np.asarray(v[key].data, np.float64))
                    else:  # variable not defined on this block: NaN marks it
                        parts.append(np.full((nt, self.block_sizes[b - 1]), np.nan))
                self.elem[name] = np.concatenate(parts, axis=1)


def read(path) -> ExodusFile:
    return ExodusFile(path)


def problem_from_exodus(path, phys, bcs, datum_elem: int = 0, datum_value: float = 0.0,
                        name: str | None = None, meta: dict | None = None,
                        renumber: bool = False) -> Problem:
    """A Problem on the mesh of an Exodus II file.

    ``bcs`` lists Dirichlet conditions as Goma's input deck does
    (``BC = U NS 1 0.`` ...): tuples (node-set ID, variable type, value) with
    Goma's variable numbering (U1=0, U2=1, U3=2, T=3).  They are applied in
    order; where two conditions fix the same unknown of a node, the later one
    holds.  The pressure datum pins the first P1 coefficient of element
    ``datum_elem`` (numbered as in the file).  The mesh has no structured
    hierarchy (``mesh.n == 0``), so the geometric multigrid preconditioner
    does not apply to it.

    ``renumber`` orders nodes and elements so that the Jacobian's band is
    narrow (``mesh.band_order``), as the banded LU needs; ``write_solution`` then maps results back to the file's
    numbering."""
    f = read(path)
    file_mesh, file_sets = f.mesh, f.node_sets
    if not 0 <= datum_elem < file_mesh.num_elems:
        raise ValueError(f"datum element {datum_elem} is not in the mesh")
    meta = dict(meta or {})
    if renumber:
        no, eo = band_order(file_mesh)
        new_id = np.empty_like(no)
        new_id[no] = np.arange(no.size)
        f.mesh = renumber_mesh(file_mesh, no, eo)
        f.node_sets = {k: np.sort(new_id[v]) for k, v in file_sets.items()}
        datum_elem = int(np.flatnonzero(eo == datum_elem)[0])
        meta.update(exodus_node_order=no, exodus_elem_order=eo, exodus_file_mesh=file_mesh,
                    exodus_node_sets=file_sets)
    bc = _BC()
    for ns, var, value in bcs:
        if ns not in f.node_sets:
            raise ValueError(f"node set {ns} is not in {path} (it has {sorted(f.node_sets)})")
        if var == PRES:
            raise ValueError("pressure is a P1 element field here; use the pressure datum")
        if var not in range(f.mesh.dim) and var != TEMP:
            raise ValueError(f"variable type {var} is not on this path")
        if var == TEMP and not phys.energy:
            raise ValueError("a temperature condition needs the energy equation")
        bc.add(f.node_sets[ns], var, value)
    nn, vv, xx = bc.arrays()
    return Problem(name=name or f"exodus_{f.title or 'mesh'}", mesh=f.mesh, phys=phys, dbc_node=nn,
                   dbc_var=vv, dbc_value=xx, datum_elem=datum_elem, datum_value=datum_value,
                   meta=meta)
