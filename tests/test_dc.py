"""The package's .dc version-4 reader and writer (pamhd_amd/dc.py), against
the reference's golden result files (tests/mhd/*/ goldens held under
tests/golden/dc) and by round trips."""
import os

import numpy as np
import pytest

import golden_cases as G
import oracle_lib as O
from pamhd_amd import dc


@pytest.mark.parametrize("case", sorted(G.GOLDENS))
def test_reader_matches_golden_files(case):
    path = os.path.join(G.DC_DIR, case + ".dc")
    meta, ids, data = dc.read(path)
    m2, ids2, mhd2 = G.read_dc(path)
    assert meta["sim_step"] == m2["sim_step"] and meta["sim_time"] == m2["sim_time"]
    np.testing.assert_array_equal(ids, ids2)
    np.testing.assert_array_equal(data["mhd     "], mhd2)


@pytest.mark.parametrize("case", ["density_advection_1d_px", "alfven_2d_x_bgB1"])
def test_written_result_reads_back_and_meets_the_golden(case, tmp_path):
    """The oracle's final state of a golden case written with dc.write and
    read back passes the reference's mhd_diff rule against the golden file,
    and every header field written round-trips."""
    cfg = G.GOLDENS[case]
    cells, length, state, bg = G.initial_state(cfg)
    g = O.make_grid(cells, length, cfg["gamma"], cfg["mu0"], bg)
    st, steps, t = O.run_periodic(g, O.RUSANOV, cfg["cfl"], 0.0, cfg["t_end"], state,
                                  max_steps=100000)
    assert st == 0
    out = tmp_path / "mhd.dc"
    gold_meta, gold_ids, gold_mhd = G.read_dc(os.path.join(G.DC_DIR, case + ".dc"))
    dc.write(out, state, cells, gold_meta["grid_start"], length, gold_meta["periodicity"],
             sim_step=steps, sim_time=t, adiabatic_index=cfg["gamma"],
             proton_mass=cfg["mp"], vacuum_permeability=cfg["mu0"])
    meta, ids, data = dc.read(out)
    for k in ("grid_start", "periodicity", "sim_time", "sim_step", "adiabatic_index",
              "proton_mass", "vacuum_permeability"):
        assert tuple(np.atleast_1d(meta[k])) == tuple(np.atleast_1d(gold_meta[k])), k
    assert tuple(meta["ref_lvl_0_cells"]) == tuple(cells)
    mine = dc.to_grid_order(ids, data["mhd     "], cells)
    gold = G.dc_to_grid(gold_meta, gold_ids, gold_mhd)
    for k, (viol, _) in G.mhd_diff(mine, gold).items():
        assert viol <= 1e-15, k
    # face copies: the -x face of each cell is the stored face, +x its neighbour's
    faces = dc.to_grid_order(ids, data["faceB   "], cells)
    np.testing.assert_array_equal(faces[0], state[5])
    np.testing.assert_array_equal(faces[2], state[6])
    np.testing.assert_array_equal(faces[4], state[7])


def test_box_state_round_trip(tmp_path):
    from pamhd_amd import configs
    sb = configs.solar_wind_box(10, 5e6, 5e5, 3e-9, -4e-9)
    b = O.Box(sb)
    s = b.init()
    out = tmp_path / "box.dc"
    dc.write(out, s, sb.cells, sb.start, sb.cell_length, (0, 0, 0), sim_step=3, sim_time=1.5,
             adiabatic_index=sb.adiabatic_index, proton_mass=sb.proton_mass,
             vacuum_permeability=sb.vacuum_permeability, cell_type=b.ctype)
    meta, ids, data = dc.read(out)
    np.testing.assert_array_equal(dc.to_grid_order(ids, data["mhd     "], sb.cells), s[:8])
    np.testing.assert_array_equal(dc.to_grid_order(ids, data["faceB   "], sb.cells), s[8:])
    np.testing.assert_array_equal(dc.to_grid_order(ids, data["mhd info"], sb.cells)[0], b.ctype)
