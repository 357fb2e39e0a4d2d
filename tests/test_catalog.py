"""Catalog parquet I/O in the reference's version-2 row format (src/jax_fli/io/catalog.py): the
schema (column names and feature types of _build_features, catalog.py:136-178), exact round trips
of every field type through a datasets.Dataset and a parquet file, batched fields, the serving
results' catalog, and the refusal of massive neutrinos."""

from __future__ import annotations

import numpy as np
import pyarrow.parquet as pq
import pytest
import torch

CPU = torch.device("cpu")


@pytest.fixture(scope="module")
def io(pkg):
    return pkg.io


def _fields(pkg):
    rng = np.random.default_rng(0)
    t = lambda *s: torch.from_numpy(rng.standard_normal(s).astype(np.float32))
    common = dict(mesh_size=(8, 8, 8), box_size=(100.0, 100.0, 100.0))
    dyn1 = dict(z_sources=0.5, scale_factors=0.66, comoving_centers=120.0, density_width=30.0)
    dyn3 = dict(z_sources=np.array([0.1, 0.2, 0.3]), scale_factors=np.array([0.9, 0.8, 0.7]),
                comoving_centers=np.array([10.0, 20.0, 30.0]), density_width=np.array([5.0, 5.0, 5.0]))
    return [
        pkg.DensityField(array=t(8, 8, 8), status=pkg.FieldStatus.DENSITY_FIELD, unit=pkg.DensityUnit.DENSITY,
                         **common, **dyn1),
        pkg.DensityField(array=t(3, 8, 8, 8), status=pkg.FieldStatus.DENSITY_FIELD, unit=pkg.DensityUnit.DENSITY,
                         **common, **dyn3),
        pkg.ParticleField(array=t(4, 4, 4, 3), status=pkg.FieldStatus.PARTICLES,
                          unit=pkg.PositionUnit.GRID_RELATIVE, **common, **dyn1),
        pkg.SphericalDensity(array=t(3, 12 * 4**2), nside=4, status=pkg.FieldStatus.LIGHTCONE,
                             unit=pkg.DensityUnit.DENSITY, observer_position=(0.5, 0.5, 0.0), **common, **dyn3),
        pkg.SphericalKappaField(array=t(12 * 2**2), nside=2, status=pkg.FieldStatus.KAPPA,
                                unit=pkg.ConvergenceUnit.DIMENSIONLESS, **common, **dyn1),
        pkg.FlatDensity(array=t(6, 5), flatsky_npix=(6, 5), field_size=(2.0, 2.0), status=pkg.FieldStatus.LIGHTCONE,
                        unit=pkg.DensityUnit.DENSITY, **common, **dyn1),
    ]


def _same(a, b):
    assert type(a) is type(b)
    assert torch.equal(a.array.cpu(), b.array.cpu())
    for k in ("mesh_size", "box_size", "observer_position", "halo_size", "nside", "flatsky_npix", "status", "unit"):
        assert getattr(a, k) == getattr(b, k), k
    for k in ("z_sources", "scale_factors", "comoving_centers", "density_width"):
        assert np.allclose(np.asarray(getattr(a, k), dtype=np.float64), np.asarray(getattr(b, k)), rtol=0, atol=0), k


def test_schema_matches_the_reference_v2_format(pkg, io):
    f = _fields(pkg)[3]
    ds = io.Catalog(f, pkg.Cosmology()).to_dataset()
    feat = ds.features
    expect = {"array", "z_sources", "scale_factors", "comoving_centers", "density_width", "mesh_size", "box_size",
              "observer_position", "halo_size", "status", "unit", "Omega_c", "Omega_b", "h", "n_s", "sigma8", "w0",
              "wa", "Omega_k", "Omega_nu", "version", "field_type", "nside"}
    assert set(feat) == expect
    assert type(feat["array"]).__name__ == "Array2D" and feat["array"].shape == (None, 12 * 16)
    assert feat["z_sources"].feature.dtype == "float64"
    assert feat["mesh_size"].length == 3 and feat["mesh_size"].feature.dtype == "int32"
    assert feat["box_size"].feature.dtype == "float32" and feat["halo_size"].length == 2
    assert feat["Omega_c"].dtype == "float32" and feat["version"].dtype == "int32"
    assert ds[0]["field_type"] == "SphericalDensity" and ds[0]["version"] == io.CATALOG_VERSION
    assert ds[0]["status"] == "LIGHTCONE" and ds[0]["unit"] == "DENSITY"


def test_round_trip_every_field_type(pkg, io, tmp_path):
    fields = _fields(pkg)
    # one catalog per field type (entries of a catalog share one schema, as in the reference), and one
    # with a batched and an unbatched density entry
    groups = [[fields[0], fields[1]]] + [[f] for f in fields[2:]]
    for gi, group in enumerate(groups):
        cosmos = [pkg.Cosmology(Omega_c=0.2 + 0.01 * i, sigma8=0.75 + 0.01 * i) for i in range(len(group))]
        cat = io.Catalog(group, cosmos)
        back = io.Catalog.from_dataset(cat.to_dataset(), device=CPU)
        path = str(tmp_path / f"cat{gi}.parquet")
        cat.to_parquet(path)
        back2 = io.Catalog.from_parquet(path, device=CPU)
        assert len(back) == len(back2) == len(group)
        for src, a, b in zip(group, back.field, back2.field):
            _same(src, a)
            _same(src, b)
        for c, a in zip(cosmos, back2.cosmology):
            assert a.Omega_c == pytest.approx(c.Omega_c, rel=1e-6) and a.sigma8 == pytest.approx(c.sigma8, rel=1e-6)
        table = pq.read_table(path)
        assert table.num_rows == len(group) and "Omega_nu" in table.column_names


def test_slicing_and_validation(pkg, io):
    fields = _fields(pkg)[:2]
    cat = io.Catalog(fields, [pkg.Cosmology(), pkg.Cosmology(Omega_c=0.3)])
    assert len(cat[1:]) == 1 and cat[1].cosmology[0].Omega_c == 0.3
    with pytest.raises(ValueError):
        io.Catalog(fields, [pkg.Cosmology()])
    ds = io.Catalog(fields[0], pkg.Cosmology()).to_dataset()
    row = dict(ds.with_format("numpy")[0])
    row["Omega_nu"] = 0.1
    with pytest.raises(ValueError):
        io.Catalog.from_dataset(row, device=CPU)


def test_serving_results_to_catalog(pkg, io):
    S = pkg.serving
    reqs = [S.SimRequest(rid=i, mesh=8, seed=i, Omega_c=0.25 + 0.01 * i, sigma8=0.8) for i in range(2)]
    res = [S.SimResult(rid=r.rid, mesh=8, displacement=torch.zeros(8, 8, 8, 3), density=torch.ones(8, 8, 8) * r.rid,
                       arrival=0.0, admitted=0.0, finished=1.0) for r in reqs]
    cat = io.catalog_from_results(res, reqs)
    back = io.Catalog.from_dataset(cat.to_dataset(), device=CPU)
    assert [float(f.array.mean()) for f in back.field] == [0.0, 1.0]
    assert back.field[0].box_size == (31.2, 31.2, 31.2) or back.field[0].box_size[0] == pytest.approx(31.2)
    assert back.cosmology[1].Omega_c == pytest.approx(0.26, rel=1e-6)
