"""The reference's dump and restart layouts (output.F90:184-476,
setup_flow.F90:448-552, setup_domain.F90:206-242): a file written the way the
reference's list-directed WRITE lays it out is read back field by field, and
the node dump round-trips through the nodes-restart reader with its index and
type checks."""
import numpy as np
import pytest

from sunset_amd import io
from sunset_amd.nodes import jittered_lattice


def test_read_fields_in_the_reference_layout(tmp_path):
    # header and rows as output_layer writes them (Fortran list-directed, E and D exponents)
    rows = [
        "   1.0000000000000000E-002  0.50000000000000000      -0.25000000000000000"
        "        3.0000000000000000        4.0000000000000000        0.0000000000000000"
        "        1.5000000000000000      -0.20000000000000001       0.90000000000000002"
        "        1.0000000000000000     ",
        "  -2.0000000000000000D-003  0.0000000000000000        1.0000000000000000"
        "       0.0000000000000000        0.0000000000000000        1.0000000000000000E-005"
        "        1.0000000000000000        0.0000000000000000        1.0000000000000000"
        "        1.0000000000000000     ",
    ]
    text = "\n".join([" !! header !!", "           2",
                      "   1.0000000000000000        0.0000000000000000        0.0000000000000000"
                      "        0.0000000000000000        0.0000000000000000     ",
                      "   1.0000000000000000E-005   2.0000000000000002E-005   3.0000000000000001E-005"
                      "   1.0000000000000000E-004",
                      " !! end header !!"] + rows) + "\n"
    p = tmp_path / "fields_10000"
    p.write_text(text)
    R = io.read_fields(str(p))
    assert R["npfb"] == 2
    np.testing.assert_array_equal(R["ro"], [1.01, 1.0 - 2e-3])
    np.testing.assert_array_equal(R["u"], [0.5, 0.0])
    np.testing.assert_array_equal(R["v"], [-0.25, 1.0])
    np.testing.assert_array_equal(R["cxx"], [1.5, 1.0])
    np.testing.assert_array_equal(R["cxy"], [-0.2, 0.0])
    np.testing.assert_array_equal(R["cyy"], [0.9, 1.0])
    assert (R["emax_np1"], R["emax_n"], R["emax_nm1"], R["dt"]) == (1e-5, 2e-5, 3e-5, 1e-4)
    assert R["eflow_nm1"] == 1.0 and R["sum_eflow"] == 0.0


def test_nodes_dump_round_trips_through_the_restart_reader(tmp_path):
    ns = jittered_lattice(12, seed=3, h_over_s=2.4)
    ns.h = ns.h * (1.0 + 0.01 * np.arange(ns.npfb) / ns.npfb)
    ns.h_small = ns.h * 0.9
    p = str(tmp_path / "nodes_10000")
    io.write_nodes(p, ns)
    base = jittered_lattice(12, seed=3, h_over_s=2.4)
    back = io.read_nodes_restart(p, base)
    np.testing.assert_array_equal(back.h, ns.h)
    np.testing.assert_array_equal(back.h_small, ns.h_small)
    np.testing.assert_array_equal(back.x, base.x)
    bad = jittered_lattice(12, seed=3, h_over_s=2.4)
    bad.node_type = bad.node_type.copy()
    bad.node_type[5] = 0
    with pytest.raises(ValueError):
        io.read_nodes_restart(p, bad)


def test_ipart_round_trip_with_boundaries(tmp_path):
    """write_ipart writes boundary nodes with their normals and leaves the FD
    rows out; read_ipart regenerates the rows along the normal
    (setup_domain.F90:179-193) and gives the same node set, bit for bit."""
    from sunset_amd.nodes import inflow_outflow_channel, read_ipart, wall_channel, write_ipart
    for ns in (wall_channel(16, 16, order=6, seed=4), inflow_outflow_channel(16, 16, order=4,
                                                                              seed=5)):
        p = str(tmp_path / "IPART")
        write_ipart(p, ns)
        with open(p) as f:
            head = f.readline().split()
        assert int(head[0]) == ns.nb and int(head[1]) == ns.npfb - 4 * ns.nb
        order = 6 if ns.meta["kind"] == "wall_channel" else 4
        r = read_ipart(p, order=order)
        for k in ("x", "y", "s", "h", "h_small", "node_type", "rnx", "rny"):
            assert np.array_equal(getattr(r, k), getattr(ns, k)), k
        assert r.bc == ns.bc
