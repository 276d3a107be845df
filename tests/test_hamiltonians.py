"""Model construction: the reference CLI's term order and counts, and the frozen PCG streams."""
from ddevol_amd import hamiltonians as H


def test_pcg_known_outputs():
    """engine/tests/test_indexed.cpp:27-30."""
    g7, gk = H.InputPCG(7), H.InputPCG(20260711)
    assert g7.next() == 0xa00641a9f1e54a8b and g7.next() == 0xe5afcdbcaf266a95
    assert gk.next() == 0x32dde8bfd7bf31ca and gk.next() == 0xe7e784102d982562


def test_edge_order_and_skip():
    t = H.edge([], 2, 5, 1.0)
    assert t == [(2, 5, 3, 3, 0.5), (2, 5, 1, 1, -0.25), (2, 5, 2, 2, -0.25)]
    assert H.edge([], 0, 1, 0.0) == []  # h == 0 terms are skipped (cli_rev.c:108-111)


def test_config_term_counts():
    assert len(H.ring(12)) == 36
    t16 = H.j1j2_n16_terms(16)
    assert len(t16) == 96
    assert t16[3][:2] == (0, 2) and t16[6][:2] == (1, 2)
    j2 = H.j1j2_n16_couplings()
    assert all(0.1 <= x <= 0.4 for x in j2) and len(set(j2)) == 16
    assert len(H.cluster_pairs()) == 158  # 54 + 72 + 32 pairs with r^2 in {1, 2, 3}
    o = H.powder_orientations(512)
    assert len(o) == 512 and all(abs(x * x + y * y + z * z - 1) < 1e-12 for x, y, z in o)
    assert len(H.range_ring(64, 4)) == 768
    assert len(H.cubic(3)) == 3 * 243  # 27 sites x (6 + 12) neighbours / 2, three terms each
