"""Host-side mirror of the reference interface: catalog construction and naming,
input validation and the reference's error classes, grid flatten / restore
(InterpolationCatalog.jl:71-225, InterpolationInput.jl:85-118,
LineSamples_interpolation.jl:48-52, transform.jl:53-116).  No GPU needed."""
import numpy as np
import pytest

from distorteddustydiscanalysistoolkits_amd import partia as P


def test_catalog_rho_v_field_set():
    cat = P.InterpolationCatalog(("rho", "vx", "vy", "vz"), scalars=("rho",), gradients=("rho",),
                                 divergences=("v",), curls=("v",))
    assert cat.ordered_names == ("rho", "∇rhoˣ", "∇rhoʸ", "∇rhoᶻ", "∇⋅v", "∇×vˣ", "∇×vʸ", "∇×vᶻ")
    assert cat.L == 8
    assert cat.scalar_snormalization == (False,)
    assert cat.scalar_slots == (0,) and cat.grad_slots == (0,)
    assert cat.div_slots == ((1, 2, 3),) and cat.curl_slots == ((1, 2, 3),)


def test_catalog_shepard_flags_and_errors():
    cat = P.InterpolationCatalog(("ρ", "dens", "T"), scalars=("ρ", "dens", "T"))
    assert cat.scalar_snormalization == (False, True, True)
    with pytest.raises(P.ArgumentError):
        P.InterpolationCatalog(("a", "a"), scalars=("a",))
    with pytest.raises(KeyError):
        P.InterpolationCatalog(("rho",), gradients=("ρ",))    # README example defect (P/README.md:224-238)
    with pytest.raises(P.DimensionMismatch):
        P.InterpolationCatalog.from_slots(scalar_names=("a", "b"), scalar_slots=(0,))
    c = P.InterpolationCatalog.from_slots(grad_names=("rho",), grad_slots=(-1,))
    assert c.ordered_names == ("∇rhoˣ", "∇rhoʸ", "∇rhoᶻ")


def test_input_validation():
    a = np.zeros(4)
    with pytest.raises(P.DimensionMismatch):
        P.InterpolationInput(a, a, np.zeros(3), a, a, a)
    with pytest.raises(P.DimensionMismatch):
        P.InterpolationInput(a, a, a, a, a, a, (np.zeros(5),))
    with pytest.raises(TypeError):
        P.InterpolationInput(a, a, a, a, a, a.astype(np.float32))
    inp = P.InterpolationInput(a, a, a, a, a, a, (a,), smoothed_kernel=P.C4_Wendland)
    assert len(inp) == 4 and inp.smoothed_kernel is P.C4_Wendland


def test_line_samples_rejects_non_scatter():
    a = np.zeros(4)
    inp = P.InterpolationInput(a, a, a, a, a, a, (a,))
    cat = P.InterpolationCatalog(("q",), scalars=("q",))
    lines = P.LineSamples(np.zeros(1), (np.zeros(1),) * 3, (np.zeros(1), np.zeros(1), np.ones(1)))
    for strat in (P.itpGather, P.itpSymmetric):
        with pytest.raises(P.ArgumentError):
            P.LineSamples_interpolation(P.MI355XComputeBackend(), lines, inp, cat, strat)


def test_flatten_restore_roundtrip():
    g = P.StructuredGrid.build(P.Cylindrical, (1.0, 5.0, 4), (0.0, 2 * np.pi, 6), (-1.0, 1.0, 3))
    flat = P.flatten(P.Cylindrical, g)
    assert len(flat) == 4 * 6 * 3
    # column-major: s fastest
    assert np.isclose(flat.coor[0][1] ** 2 + flat.coor[1][1] ** 2, g.axes[0][1] ** 2)
    flat.grid[:] = np.arange(len(flat))
    back = P.restore_struct(P.Cylindrical, flat, g.axes)
    assert back.grid.shape == (4, 6, 3) and back.grid[1, 0, 0] == 1.0 and back.grid[0, 1, 0] == 4.0
    with pytest.raises(P.ArgumentError):
        P.restore_struct(P.Cartesian, flat, g.axes)
    c = P.StructuredGrid.build(P.Cartesian, (0.2, 0.7, 3), (0.25, 0.55, 2), (0.3, 0.6, 2))
    fc = P.flatten(P.Cartesian, c)
    assert np.allclose(fc.coor[0][:3], [0.2, 0.45, 0.7], rtol=0, atol=1e-16) and fc.coor[1][3] == 0.55


def test_spherical_grid_axes_flatten_and_errors():
    """StructuredGrid(Spherical, ...) (StructuredGrid.jl:427-518): r on a LinRange (or one shell
    radius), phi half-open, theta = acos of a LinRange in cos(theta); flatten in column-major order
    with _sph2cart (coordinate_transformations.jl:197-203); the reference's ArgumentError texts."""
    g = P.StructuredGrid.build(P.Spherical, (1.0, 10.0, 4), (0.0, 2 * np.pi, 6), (0.0, np.pi, 5))
    ra, pa, ta = g.axes
    assert g.size == (4, 6, 5)
    assert np.array_equal(ra, [1.0, 4.0, 7.0, 10.0])
    assert np.allclose(pa, np.arange(6) * (2 * np.pi / 6), rtol=0, atol=1e-15)
    assert np.allclose(np.cos(ta), np.linspace(1.0, -1.0, 5), rtol=0, atol=1e-15)   # equal-area in cos(theta)
    flat = P.flatten(P.Spherical, g)
    x, y, z = flat.coor
    i = 0
    for k in range(5):                    # theta slowest, r fastest
        for j in range(6):
            for m in range(4):
                r, ph, th = ra[m], pa[j], ta[k]
                assert (x[i], y[i], z[i]) == (r * np.sin(th) * np.cos(ph), r * np.sin(th) * np.sin(ph), r * np.cos(th))
                i += 1
    assert np.allclose(np.sqrt(x * x + y * y + z * z), np.tile(ra, 30), rtol=1e-14)
    back = P.restore_struct(P.Spherical, flat, g.axes)
    assert back.size == g.size
    shell = P.StructuredGrid.build(P.Spherical, 5.0, (0.0, np.pi, 4), (0.1, 3.0, 3))
    assert shell.size == (1, 4, 3) and np.array_equal(shell.axes[0], [5.0])
    for bad, text in ((((1.0, 0.5, 4), (0.0, 1.0, 2), (0.0, 1.0, 2)), "radial"),
                      (((0.0, 1.0, 4), (0.0, 7.0, 2), (0.0, 1.0, 2)), "angular"),
                      (((0.0, 1.0, 4), (0.0, 1.0, 2), (0.0, 4.0, 2)), "polar"),
                      (((0.0, 1.0, 0), (0.0, 1.0, 2), (0.0, 1.0, 2)), "must be")):
        with pytest.raises(P.ArgumentError) as e:
            P.StructuredGrid.build(P.Spherical, *bad)
        assert text in str(e.value)
    with pytest.raises(P.ArgumentError):
        P.StructuredGrid.build(P.Spherical, -1.0, (0.0, 1.0, 2), (0.0, 1.0, 2))
