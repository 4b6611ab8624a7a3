"""Density summed inside the collide kernel (FP32 storage forms).

After a step, rho at every fluid node must equal the sum of the 19 populations
that the next step pulls there (MacroCalculate, FluidCuda.cu:351-372), whichever
tile, chunk or periodic image emitted them; solid nodes keep the wall values of
the step, as with the separate density pass.  Lattices that the 32 x 8 x 32
(x, y, z) tiles cover exactly take the fused path: one tile or several in each
direction (a single tile reads its own border ring through the periodic wrap),
walls of equal and unequal thickness, periodic boxes.  The other geometries
(partial tiles, ny = 1, nx below the chunk length) and lattices too small for
the fused path (single-plane chunks) take the separate density pass and are
checked the same way.  Chunks of 32 planes, the length large lattices select,
are checked in test_gpu_production.py."""
import os

import numpy as np
import pytest

import gpu_cases as gc

pytestmark = pytest.mark.gpu

GEOMETRIES = {
    "ref_prw_walls": None,
    "prw_2d": None,
    "prm_walls_2_4": None,
    "vdw_variants": None,
    # 33 x planes (an odd count); y and z tiles both partial; periodic
    "rke_33x13x21": dict(dims=(33, 13, 21), eos="RKE", tr=0.60, nwall=0,
                         spheres=[(16, 6, 10, 5)], width=4.0),
    # nx below the chunk length, two z tiles, walls 3 + 3
    "prw_5x9x40": dict(dims=(5, 9, 40), eos="PRW", tr=0.70, nwall=3,
                       spheres=[(2, 4, 3, 6)], width=4.0),
    # 70 x planes, y tiles 8 + 8 + 1, z tiles 32 + 13
    "cse_70x17x45": dict(dims=(70, 17, 45), eos="CSE", tr=0.85, nwall=0,
                         spheres=[(35, 8, 22, 9)], width=2.5, Ka=0.01),
    # one x plane; the lattice folds onto itself in x
    "prw_1x16x36": dict(dims=(1, 16, 36), eos="PRW", tr=0.70, nwall=3,
                        spheres=[(0, 8, 3, 6)], width=4.0),
}
TILED = {
    # 2 chunks x 2 y tiles x 2 z tiles, walls 3 + 3
    "tiled_prw_64x16x64": dict(dims=(64, 16, 64), eos="PRW", tr=0.70, nwall=3,
                               spheres=[(32, 8, 3, 12)], width=4.0),
    # one tile in every direction, periodic
    "tiled_cse_32x8x32": dict(dims=(32, 8, 32), eos="CSE", tr=0.85, nwall=0,
                              spheres=[(16, 4, 16, 6)], width=2.5, Ka=0.01),
    # 3 chunks, 3 y tiles, one z tile, walls 2 + 4
    "tiled_prm_96x24x32": dict(dims=(96, 24, 32), eos="PRM", tr=0.70, nwall=2, nwall_hi=4,
                               spheres=[(48, 12, 2, 9)], width=4.0),
    # one chunk, 2 y tiles, 3 z tiles, two droplets across tile borders, periodic
    "tiled_rke_32x16x96": dict(dims=(32, 16, 96), eos="RKE", tr=0.60, nwall=0,
                               spheres=[(16, 8, 31.5, 9), (16, 8, 64, 9)], width=5.0),
}
GEOMETRIES.update(TILED)


def make(cpmp, name, form, fuse=True, xc=None):
    """xc: planes per march chunk (0: the library's choice); the tiled geometries
    default to 2, which they need for the fused path (lattices this small select
    single planes and the separate density pass by themselves)"""
    spec = GEOMETRIES[name]
    old = os.environ.get("CPMP_FUSE")
    old_xc = os.environ.get("CPMP_XC")
    if not fuse:
        os.environ["CPMP_FUSE"] = "0"
    else:                      # the fused path where the geometry allows it
        os.environ.pop("CPMP_FUSE", None)
    if xc is None and (name in TILED or name in PERIODIC_TILED):
        xc = 2
    if xc:
        os.environ["CPMP_XC"] = str(xc)
    elif xc == 0:              # the library's own choice
        os.environ.pop("CPMP_XC", None)
    try:
        if spec is None:
            _, g = gc.build(gc.CASES[name], cpmp, form)
            return g
        nx, ny, nz = spec["dims"]
        p = cpmp.make_params(nx, ny, nz, eos=spec["eos"], tr=spec["tr"], nwall=spec["nwall"],
                             nwall_hi=spec.get("nwall_hi"), Ka=spec.get("Ka", 0.001), form=form)
        dg, dl = gc.densities(spec["eos"], spec["tr"])
        if form == "fp16":
            p.rho0 = dg
        g = cpmp.Lattice(p)
        g.set_wall_mu(-0.01, 0.01)
        rho = gc.sphere_rho(nx, ny, nz, spec["spheres"], spec["width"], dg, dl).reshape(nx, ny, nz)
        rho[:, :, :spec["nwall"]] = 0
        rho[:, :, nz - spec.get("nwall_hi", spec["nwall"]):nz] = 0
        g.init_from_rho(rho.ravel())
        return g
    finally:
        if old is None:
            os.environ.pop("CPMP_FUSE", None)
        else:
            os.environ["CPMP_FUSE"] = old
        if old_xc is None:
            os.environ.pop("CPMP_XC", None)
        else:
            os.environ["CPMP_XC"] = old_xc


@pytest.mark.parametrize("form", ["fp32", "fp16"])
@pytest.mark.parametrize("name", list(GEOMETRIES))
def test_rho_is_the_sum_of_pulled_populations(cpmp, name, form):
    g = make(cpmp, name, form)
    fluid = g.get("type") == 1
    for n in (1, 6):
        g.step(n)
        rho = g.get("rho")
        f = g.get("f").reshape(19, -1)
        s = f.sum(axis=0)           # FP64 sum of the values the next step pulls
        err = np.max(np.abs(rho[fluid] - s[fluid])) / np.max(np.abs(s[fluid]))
        print(f"[density] {name} {form} step {g.step_count}: max |rho - sum f| / max rho = {err:.3e}", flush=True)
        assert err < 2e-6
    g.close()


@pytest.mark.parametrize("name", list(GEOMETRIES))
def test_fused_density_matches_separate_pass(cpmp, name):
    a = make(cpmp, name, "fp32", fuse=True)
    b = make(cpmp, name, "fp32", fuse=False)
    a.step(1)
    b.step(1)
    ra, rb = a.get("rho"), b.get("rho")
    fluid = a.get("type") == 1
    # fluid: the same stored values summed in another order
    assert np.max(np.abs(ra[fluid] - rb[fluid])) <= 2e-6 * np.max(rb[fluid])
    # solid: this step's wall values in both
    assert np.array_equal(ra[~fluid], rb[~fluid])
    a.step(40)
    b.step(40)
    ra, rb = a.get("rho"), b.get("rho")
    err = np.max(np.abs(ra[fluid] - rb[fluid])) / np.max(rb[fluid])
    print(f"[density] {name}: fused vs separate after 41 steps {err:.3e}", flush=True)
    assert err < 1e-4
    assert abs(a.reduce().mass - b.reduce().mass) <= 1e-6 * b.reduce().mass
    a.close()
    b.close()


@pytest.mark.parametrize("name", list(TILED))
def test_small_tiled_lattices_select_the_separate_pass(cpmp, name):
    """left to choose, FP32 lattices too small for 2048 march workgroups of two
    planes take single planes and the separate density pass (no extra buffers);
    the density is still the sum of the pulled populations"""
    a = make(cpmp, name, "fp32", xc=0)
    b = make(cpmp, name, "fp32", fuse=False, xc=0)
    assert a.memory() == b.memory()
    fluid = a.get("type") == 1
    a.step(3)
    rho, s = a.get("rho"), a.get("f").reshape(19, -1).sum(axis=0)
    assert np.max(np.abs(rho[fluid] - s[fluid])) < 2e-6 * np.max(np.abs(s[fluid]))
    a.close()
    b.close()


@pytest.mark.parametrize("name", list(GEOMETRIES))
def test_fused_path_taken_on_tiled_lattices(cpmp, name):
    """the fused contexts hold the extra density buffers exactly when the tiles
    cover the lattice"""
    if gc.path_forced():
        pytest.skip("the environment overrides the path selection")
    a = make(cpmp, name, "fp32", fuse=True)
    b = make(cpmp, name, "fp32", fuse=False)
    extra = a.memory() - b.memory()
    a.close()
    b.close()
    assert (extra > 0) == (name in TILED)


def test_tiled_layout_fields_checkpoints_and_repeatability(cpmp, tmp_path):
    """the tile-contiguous population layout is invisible at the boundary: the
    same initial state reads the same through get_field, checkpoints written by
    one layout load into the other, and the fused path repeats bit for bit"""
    name = "tiled_prw_64x16x64"
    a = make(cpmp, name, "fp32", fuse=True)
    b = make(cpmp, name, "fp32", fuse=False)
    if not gc.path_forced():
        assert a.memory() > b.memory()
    assert np.array_equal(a.get("f"), b.get("f"))
    a.step(7)
    path = str(tmp_path / "tiled.ckp")
    a.checkpoint_save(path)
    b.checkpoint_load(path)
    assert np.array_equal(a.get("f"), b.get("f"))
    assert np.array_equal(a.get("rho"), b.get("rho"))
    c = make(cpmp, name, "fp32", fuse=True)
    c.checkpoint_load(path)
    a.step(5)
    c.step(5)
    assert np.array_equal(a.get("rho"), c.get("rho"))
    assert np.array_equal(a.get("f"), c.get("f"))
    for g in (a, b, c):
        g.close()


PERIODIC_TILED = {
    # 4 chunks x 2 y tiles x 2 z tiles, periodic: several chunk ends and tile borders
    "tiled_rke_128x16x64": dict(dims=(128, 16, 64), eos="RKE", tr=0.60, nwall=0,
                                spheres=[(40, 8, 31.5, 10), (88, 8, 32, 10)], width=5.0),
}
GEOMETRIES.update(PERIODIC_TILED)


@pytest.mark.parametrize("name", ["tiled_cse_32x8x32", "tiled_rke_32x16x96", "tiled_rke_128x16x64"])
def test_completion_folded_into_mu_is_bitwise_the_same(cpmp, name):
    """periodic FP32 lattices: within a run of steps the density completion is done
    by the next step's mu kernel instead of k_density_fix; the state after every
    run (densities, populations, mu) is bit-identical to running the completion
    pass every step (CPMP_DEFER=0)"""
    a = make(cpmp, name, "fp32")
    os.environ["CPMP_DEFER"] = "0"
    try:
        b = make(cpmp, name, "fp32")
        for n in (1, 5, 12):
            os.environ["CPMP_DEFER"] = "0"
            b.step(n)
            os.environ.pop("CPMP_DEFER", None)
            a.step(n)
            for field in ("rho", "mu", "f"):
                assert np.array_equal(a.get(field), b.get(field)), (n, field)
    finally:
        os.environ.pop("CPMP_DEFER", None)
    a.close()
    b.close()


@pytest.mark.parametrize("form", ["fp64", "fp32", "fp16", "fp32a"])
@pytest.mark.parametrize("name", ["tiled_rke_128x16x64", "tiled_prw_64x16x64"])
def test_runs_split_into_calls_are_bitwise_the_same(cpmp, name, form):
    """the last step of every call runs the collide instantiation that also writes
    MV (and, on periodic FP32 lattices, completes the density with the separate pass
    instead of the next mu kernel): one call of 6 steps and calls of 1 + 2 + 3 steps
    leave the same bits, which needs every instantiation of the collide to round
    alike"""
    a = make(cpmp, name, form)
    b = make(cpmp, name, form)
    a.step(6)
    for n in (1, 2, 3):
        b.step(n)
    for field in ("rho", "mu", "f"):
        assert np.array_equal(a.get(field), b.get(field)), (form, field)
    a.close()
    b.close()


@pytest.mark.parametrize("collide", ["march"])
@pytest.mark.parametrize("form", ["fp64", "fp32", "fp16", "fp32a"])
def test_solid_populations_read_zero_after_a_step(cpmp, form, collide):
    """the lanes of solid nodes store zeros into their own slots with the fluid
    lanes of their row (whole population rows): after a step of the march collide
    a solid node's 19 values read 0, its fluid neighbours are unaffected (the
    density stays the sum of the pulled populations)"""
    old = os.environ.get("CPMP_COLLIDE")
    os.environ["CPMP_COLLIDE"] = collide
    try:
        g = make(cpmp, "tiled_prw_64x16x64", form)
    finally:
        if old is None:
            os.environ.pop("CPMP_COLLIDE", None)
        else:
            os.environ["CPMP_COLLIDE"] = old
    fluid = g.get("type") == 1
    for n in (1, 2):
        g.step(n)
        f = g.get("f").reshape(19, -1)
        assert np.all(f[:, ~fluid] == 0), (form, collide, n)
        rho, s = g.get("rho"), f.sum(axis=0)
        assert np.max(np.abs(rho[fluid] - s[fluid])) < 2e-6 * np.max(np.abs(s[fluid]))
    g.close()
