"""Disclination finding against the reference's own finder
(QtensorToOrderDirector, src/analysis_fields.cc; DefectFinder<BC>,
src/analysis/defect_finder.tpp), built from its sources by
oracle/build_ref.sh into oracle/_ref/ref_defects.

The committed fixture tests/golden/defects_ref.npz (written by
tests/golden/make_golden.py from that binary) holds two Q fields on a
14 x 12 x 10 grid: a seeded random director field (dense, with plaquettes
where several lines meet) and two straight wedge lines. Per boundary preset
it holds the reference's order, director, pierced plaquettes and paths.

What is pinned:
- the winding test: the oracle restatement (oracle/defects_ref.py) and the
  device kernel (lbm_defect_faces) against the reference's bytes;
- the order and director: the oracle (orc_order_director) and the device
  against QtensorToOrderDirector (the director up to sign);
- the graph: the edges walked by the reference's paths equal the host
  graph's edges (lbman3d_amd/defects.py), each edge walked once, and the
  plaquette centres agree. Lines without junctions give the same walks up to
  direction. Where lines meet, the reference's split into walks follows its
  std::unordered_map iteration order, which is not reproduced (the number of
  walks may differ there).
"""
import os

import numpy as np
import pytest

from lbman3d_amd import abi, defects
from oracle import defects_ref as dr

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "defects_ref.npz")
BCS = {"FullyPeriodic": abi.FullyPeriodicConfig, "Channel": abi.ChannelConfig}


def fixture():
    z = np.load(FIX, allow_pickle=False)
    dims = tuple(int(v) for v in z["dims"])
    cases = []
    for name in ("random", "lines"):
        for bc in BCS:
            k = f"{name}_{bc}"
            lens = z[f"{k}_path_len"]
            pts = z[f"{k}_path_points"]
            paths = np.split(pts, np.cumsum(lens)[:-1]) if len(lens) else []
            cases.append(dict(name=name, bc=bc, q=z[f"{name}_q"], order=z[f"{k}_order"],
                              director=z[f"{k}_director"],
                              faces=tuple(z[f"{k}_def_{a}"] for a in "xyz"), paths=paths))
    return dims, cases


def check_graph(dims, faces, paths):
    """The reference's paths against the host graph built from the same
    pierced plaquettes."""
    g = defects.DefectGraph(*dims, *faces)
    where = {tuple(np.round(g.location(f), 9)): f for f in g.adj}
    walked = []
    ref_walks = []
    for p in paths:
        ids = [where[tuple(np.round(pt, 9))] for pt in p]      # every point is a plaquette centre of the graph
        ref_walks.append(ids)
        walked += [(min(a, b), max(a, b)) for a, b in zip(ids[:-1], ids[1:])]
    assert len(walked) == len(set(walked)), "an edge walked twice"
    assert set(walked) == g.edges()
    # components without junctions: the same walk up to direction (a loop up to its start)
    mine = g.paths()
    deg = {f: len(nb) for f, nb in g.adj.items()}

    def nodes(w):   # a closed walk repeats its first plaquette at the end
        return sorted(w[:-1] if len(w) > 1 and w[0] == w[-1] else w)
    simple_ref = sorted(nodes(w) for w in ref_walks if all(deg[f] <= 2 for f in w))
    simple_mine = sorted(nodes(w) for w in mine if all(deg[f] <= 2 for f in w))
    assert simple_ref == simple_mine
    for w in ref_walks:
        if all(deg[f] <= 2 for f in w) and w[0] != w[-1]:
            m = next(p for p in mine if nodes(p) == nodes(w))
            assert m == w or m == w[::-1]
    return g


def test_fixture_restatement_and_graph():
    dims, cases = fixture()
    nx, ny, nz = dims
    for c in cases:
        print(f"case fixture {c['name']} {c['bc']}", flush=True)
        mine = dr.defect_faces(c["director"], nx, ny, nz)
        for a, b in zip(mine, c["faces"]):
            np.testing.assert_array_equal(a, b)
        g = check_graph(dims, c["faces"], c["paths"])
        if c["name"] == "lines":
            assert len(c["paths"]) == 2 and len(g.paths()) == 2


def test_fixture_oracle_order_director():
    from oracle.pyoracle import Oracle
    dims, cases = fixture()
    for c in cases:
        cfg = abi.make_config(*dims, BCS[c["bc"]], abi.Params())
        o = Oracle(cfg)
        o.set_state(q=c["q"])
        order, director = o.order_director()
        np.testing.assert_allclose(order, c["order"], rtol=0, atol=1e-12)
        dots = np.abs(np.sum(director * c["director"], axis=1))
        np.testing.assert_allclose(dots, 1.0, rtol=0, atol=1e-12)


@pytest.mark.skipif(not os.path.exists(os.path.join(HERE, "..", "oracle", "_ref", "ref_defects")),
                    reason="reference finder not built (no reference sources or HDF5/MPI headers)")
@pytest.mark.parametrize("seed", [3, 4])
def test_live_reference_finder(seed):
    from oracle.pyoracle import run_ref_defects
    dims = (14, 12, 10)
    q = abi.seed_director_field(seed, *dims, 0.5, 0.01)
    for bc in BCS:
        r = run_ref_defects(bc, dims, q)
        mine = dr.defect_faces(r["director"], *dims)
        for a, b in zip(mine, r["faces"]):
            np.testing.assert_array_equal(a, b)
        check_graph(dims, r["faces"], r["paths"])


@pytest.mark.gpu
def test_device_matches_reference_fixture():
    """The device's order, director and pierced plaquettes from the
    fixture's Q against the reference finder's, and the device's lines
    against the reference's paths."""
    from lbman3d_amd.solver import DeviceSolver
    dims, cases = fixture()
    for c in cases:
        print(f"case device defects {c['name']} {c['bc']}", flush=True)
        cfg = abi.make_config(*dims, BCS[c["bc"]], abi.Params())
        with DeviceSolver(cfg) as s:
            s.upload_state(q=c["q"])
            faces = defects.device_defect_faces(s)
            order, director = s.order_director()
        for a, b in zip(faces, c["faces"]):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_allclose(order, c["order"], rtol=0, atol=1e-12)
        np.testing.assert_allclose(np.abs(np.sum(director * c["director"], axis=1)), 1.0, rtol=0, atol=1e-12)
        check_graph(dims, faces, c["paths"])
