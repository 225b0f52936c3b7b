"""Output files and fields (exDyna3D::output_results, DataOut::write_vtu) on
the CPU: the .vtu writer, and the oracle's output snapshots pinned by
independent recomputation and by the nodal-update identities."""

import numpy as np
import pytest

from enscdynamics_amd import decks
from enscdynamics_amd import engine as eng
from enscdynamics_amd.vtu import read_vtu, write_vtu
from oracle.oracle import Oracle


def two_hexes():
    x = np.array([[i, j, k] for k in (0.0, 1.0) for j in (0.0, 1.0) for i in (0.0, 1.0, 2.0)])
    idx = lambda i, j, k: i + 3 * j + 6 * k  # noqa: E731
    conn = np.array([[idx(i, 0, 0), idx(i + 1, 0, 0), idx(i + 1, 1, 0), idx(i, 1, 0),
                      idx(i, 0, 1), idx(i + 1, 0, 1), idx(i + 1, 1, 1), idx(i, 1, 1)] for i in (0, 1)],
                    dtype=np.uint32)
    return x, conn


def test_vtu_round_trip(tmp_path):
    x, conn = two_hexes()
    rng = np.random.default_rng(0)
    pd = {"a": rng.normal(size=(12, 3)), "mass": rng.uniform(1, 2, size=12)}
    cd = {"vonMises": rng.uniform(size=2), "Stress": rng.normal(size=(2, 6))}
    p = str(tmp_path / "s.vtu")
    write_vtu(p, x, conn, 1.25e-3, pd, cd)
    r = read_vtu(p)
    assert r["n_points"] == 12 and r["n_cells"] == 2
    assert np.array_equal(r["points"], x.astype(np.float32))
    assert np.array_equal(r["connectivity"], conn.reshape(-1).astype(np.int64))
    assert np.array_equal(r["offsets"], np.array([8, 16]))
    assert np.array_equal(r["types"], np.array([12, 12], dtype=np.uint8))  # VTK_HEXAHEDRON
    assert r["time"] == np.float32(1.25e-3)
    assert list(r["point_data"]) == ["a", "mass"] and list(r["cell_data"]) == ["vonMises", "Stress"]
    assert np.array_equal(r["point_data"]["a"], pd["a"].astype(np.float32))
    assert np.array_equal(r["point_data"]["mass"], pd["mass"].astype(np.float32))
    assert np.array_equal(r["cell_data"]["Stress"], cd["Stress"].astype(np.float32))
    assert r["cell_data"]["Stress"].shape == (2, 6)
    # the XML part names the arrays the way VTK's reader expects
    head = open(p, "rb").read(2000).decode("latin-1")
    assert 'type="UnstructuredGrid"' in head and 'header_type="UInt64"' in head
    assert 'Name="Stress" NumberOfComponents="6"' in head


def test_vtu_rejects_bad_cells(tmp_path):
    x, conn = two_hexes()
    bad = conn.copy()
    bad[1, 3] = 99
    with pytest.raises(eng.EnscError):
        write_vtu(str(tmp_path / "b.vtu"), x, bad, 0.0)


@pytest.fixture(scope="module")
def c1_outputs():
    m = decks.config1()
    o = Oracle(m, 0).set_output(0.0).init()
    o.run(0, [], 0)
    return m, o


def test_oracle_output_cadence(c1_outputs):
    # calculate_time_interval: fewer than 10 frames at 50 per second, so
    # T / 10; output at t = 0, then whenever time passes the next multiple,
    # and at the total time
    m, o = c1_outputs
    T = m.total_time
    ti = T / 10.0
    n = o.snapshots()
    times = np.array([o.snapshot(k, "time")[0] for k in range(n)])
    assert times[0] == 0.0 and times[-1] == pytest.approx(T, rel=1e-14)
    assert np.all(np.diff(times) > 0)
    t_out = 0.0
    for t in times[:-1]:
        assert t >= t_out
        t_out += ti
    assert n in (11, 12)


def test_oracle_output_fields_recomputed(c1_outputs):
    m, o = c1_outputs
    B0 = o.get("B0")  # [Ne][8][3]
    conn = m.conn.astype(np.int64)
    for k in (1, o.snapshots() // 2):
        u = o.snapshot(k, "u")
        st = o.snapshot(k, "Stress")
        # von Mises as sqrt(3 J2) of the deviator (xy, yz, xz from the VTK order)
        s = np.zeros((m.n_elems, 3, 3))
        s[:, 0, 0], s[:, 1, 1], s[:, 2, 2] = st[:, 0], st[:, 1], st[:, 2]
        s[:, 0, 1] = s[:, 1, 0] = st[:, 3]
        s[:, 1, 2] = s[:, 2, 1] = st[:, 4]
        s[:, 0, 2] = s[:, 2, 0] = st[:, 5]
        dev = s - np.trace(s, axis1=1, axis2=2)[:, None, None] / 3 * np.eye(3)
        vm = np.sqrt(1.5 * np.einsum("eij,eij->e", dev, dev))
        assert np.allclose(o.snapshot(k, "vonMises"), vm, rtol=1e-11, atol=1e-9 * vm.max())
        # Green-Lagrange strain from F = I + sum_k u_k (x) B0_k
        F = np.eye(3) + np.einsum("eki,ekj->eij", u[conn], B0)
        E = 0.5 * (np.einsum("eki,ekj->eij", F, F) - np.eye(3))
        gl = o.snapshot(k, "Green-Lagrange Strain")
        ref = np.stack([E[:, 0, 0], E[:, 1, 1], E[:, 2, 2], E[:, 0, 1], E[:, 1, 2], E[:, 0, 2]], axis=1)
        assert np.allclose(gl, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


def test_oracle_output_is_the_state_before_the_update(c1_outputs):
    # the last snapshot is taken in the last step after compute_a: the final
    # state is that snapshot pushed through update_velocity and
    # update_displacement (exDyna3D.cpp:646-654), and the final stress is
    # the snapshot's stress
    m, o = c1_outputs
    k = o.snapshots() - 1
    sc = o.scalars()
    hs = sc["dt_i"] + sc["dt_i_1"]
    v_end = o.snapshot(k, "v") + hs * o.snapshot(k, "a") / 2.0
    assert np.array_equal(v_end, o.get("v"))
    assert np.array_equal(o.snapshot(k, "u") + sc["dt_i"] * v_end, o.get("u"))
    sg = o.get("sigma")
    assert np.array_equal(o.snapshot(k, "Stress"), sg[:, [0, 1, 2, 3, 5, 4]])


def test_history_csv_layout(tmp_path):
    rows = np.arange(2 * (len(eng.H_NAMES) + 6), dtype=np.float64).reshape(2, -1) * 1e-3
    p = str(tmp_path / "h.csv")
    assert eng.write_history_csv(p, rows, [7, 9]) == 2
    lines = open(p).read().splitlines()
    assert lines[0].split(",")[:3] == ["step", "time", "dt_i"]
    assert lines[0].split(",")[-3:] == ["ux_node9", "uy_node9", "uz_node9"]
    back = np.array([[float(v) for v in ln.split(",")[1:]] for ln in lines[1:]])
    assert np.array_equal(back, rows)
