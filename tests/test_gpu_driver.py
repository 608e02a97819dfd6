"""The drop-in driver: the gpu-vah executable with the reference's command
line and property files, its .dat output (rhic/src/FileIO.cpp:14-59) and its
cadence (rhic/src/HydroPlugin.cpp:208-233), checked against the oracle run
from the same initial condition."""
import os
import subprocess

import numpy as np
import pytest

from tests.parity import vah

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gpu-vah_amd", "gpu-vah")


def write_conf(d, nx=21, ny=17, nt=20, ictype=2):
    (d / "lattice.properties").write_text(
        f"numLatticePointsX={nx}\nnumLatticePointsY={ny}\nnumLatticePointsRapidity=1\nnumProperTimePoints={nt}\n"
        "latticeSpacingX=0.3\nlatticeSpacingY=0.3\nlatticeSpacingRapidity=0.1\nlatticeSpacingProperTime=0.01\n")
    (d / "ic.properties").write_text(
        f"initialConditionType={ictype}\nnumberOfNucleonsPerNuclei=197\ninitialEnergyDensity=15.6269\n"
        "scatteringCrossSectionNN=42.0\nimpactParameter=7.0\nfractionOfBinaryCollisions=0.5\n")
    (d / "hydro.properties").write_text(
        "initialProperTimePoint=0.5\nshearViscosityToEntropyDensity=0.2\nfreezeoutTemperatureGeV=0.140\n"
        "initializePimunuNavierStokes=0\n")


def test_cli_run_writes_reference_files(tmp_path, oracle_mod):
    conf, out = tmp_path / "conf", tmp_path / "out"
    conf.mkdir()
    out.mkdir()
    write_conf(conf)
    r = subprocess.run([EXE, "--config", str(conf), "-o", str(out), "-h", "-b"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Done hydro." in r.stdout
    assert r.stdout.count("Elapsed time/step") == 3  # n-1 = 0, 10, 20 of nt+1 = 21 steps
    names = ["e", "p", "ux", "uy", "un", "ut", "pl", "pixx", "pixy", "pixn", "piyy", "piyn", "pitt", "pitx",
             "pity", "pitn", "pinn"]
    for tau in ("0.500", "0.600", "0.700"):
        for nm in names:
            assert (out / f"{nm}_{tau}.dat").exists(), (nm, tau)
    # FileIO.cpp:46-56 format: x y z value, z outer, x inner, %.3f %.3f %.3f %.8f
    lines = (out / "e_0.700.dat").read_text().splitlines()
    assert len(lines) == 21 * 17
    x0, y0, z0, _ = lines[0].split("\t")
    assert (x0, y0, z0) == ("-3.000", "-2.400", "0.000")
    assert lines[1].split("\t")[0] == "-2.700"

    # the same run on the oracle from the same reference initial condition
    V = vah()
    c = V.config_load(str(conf))
    e, ut, ux, uy, un = V.ic_reference(c)
    o = oracle_mod.Oracle(21, 17, 1, 0.3, 0.3, 0.1, 0.01, 0.2, 0.5)
    o.set_primitives(e, ut, ux, uy, un)  # P_L = 0.0154483 e as the reference initialises it
    o.step(20)
    st = o.state()
    ebin = np.fromfile(out / "e_0.700.bin", dtype=np.float64)
    assert np.max(np.abs(ebin - st["e"]) / st["e"]) < 1e-10
    edat = np.array([float(l.split("\t")[3]) for l in lines])
    assert np.max(np.abs(edat - st["e"])) <= 5.1e-9  # %.8f rounding


def test_plugin_run_symbol_matches_cli(tmp_path):
    """The reference's run() symbol, called by a program shaped like the
    reference's Run.cpp with the reference's parameter structs
    (tests/plugin_caller.cpp), writes the same files, byte for byte, as the
    gpu-vah command line given the same properties."""
    from tests.test_plugin_abi import build_caller
    exe = build_caller(tmp_path)
    conf, out_cli, out_plugin = tmp_path / "conf", tmp_path / "cli", tmp_path / "plugin"
    for d in (conf, out_cli, out_plugin):
        d.mkdir()
    write_conf(conf, nx=21, ny=17, nt=20, ictype=2)
    r = subprocess.run([EXE, "--config", str(conf), "-o", str(out_cli), "-h"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe, str(out_plugin), "21", "17", "20", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "hydro failed" not in r.stderr, r.stderr
    assert r.stdout.count("Elapsed time/step") == 3
    files = sorted(p.name for p in out_cli.iterdir())
    assert len(files) == 17 * 3 and files == sorted(p.name for p in out_plugin.iterdir())
    for name in files:
        assert (out_cli / name).read_bytes() == (out_plugin / name).read_bytes(), name
