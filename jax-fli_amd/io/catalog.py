"""Catalog: fields + cosmologies with HuggingFace ``datasets`` / parquet serialization, in the
reference's version-2 row format (src/jax_fli/io/catalog.py:28-510).

One row per (field, cosmology) entry. The row holds the whole batched array (leading batch
dimension B, added when the field is unbatched and squeezed back on load when B == 1), the dynamic
metadata (z_sources, scale_factors, comoving_centers, density_width) as float64 sequences of length
B, the static metadata (mesh_size, box_size, observer_position, halo_size, status and unit names,
optional nside / flatsky_npix / field_size), the nine cosmological parameters as float32, the
version and the field type name -- the column names and feature types of ``_build_features``
(catalog.py:113-178), so files written here and by the reference read the same way.

Arrays come off the device once per entry (host float32); fields read back are placed on
``device`` (default: the current GPU, else the CPU). Massive neutrinos are not modelled by this
engine's cosmology: ``Omega_nu`` is written as 0 and a nonzero value is refused on load.
"""

from __future__ import annotations

import numpy as np
import torch

from ..cosmology import Cosmology
from ..fields import (AbstractField, ConvergenceUnit, DensityField, DensityUnit, FieldStatus, FlatDensity,
                      FlatKappaField, ParticleField, PositionUnit, SphericalDensity, SphericalKappaField)

CATALOG_VERSION = 2

_CLASSES = {
    "SphericalDensity": SphericalDensity,
    "SphericalKappaField": SphericalKappaField,
    "FlatDensity": FlatDensity,
    "FlatKappaField": FlatKappaField,
    "DensityField": DensityField,
    "ParticleField": ParticleField,
}
# array rank without the batch dimension (catalog.py:56-85)
_BASE_NDIM = {"SphericalDensity": 1, "SphericalKappaField": 1, "FlatDensity": 2, "FlatKappaField": 2,
              "DensityField": 3, "ParticleField": 4}
_COSMO = ("Omega_c", "Omega_b", "h", "n_s", "sigma8", "w0", "wa", "Omega_k", "Omega_nu")


def _host_array(field: AbstractField) -> np.ndarray:
    name = type(field).__name__
    if name not in _BASE_NDIM:
        raise ValueError(f"Unknown field type: {name}")
    arr = field.array
    arr = arr.detach().cpu().numpy() if isinstance(arr, torch.Tensor) else np.asarray(arr)
    nd = _BASE_NDIM[name]
    if arr.ndim == nd:
        arr = arr[None]
    elif arr.ndim != nd + 1:
        raise ValueError(f"Unexpected shape for {name}: {arr.shape}")
    return arr


def _dynamic(value, name: str, batch: int) -> list:
    """One float per batch element (catalog.py:88-110)."""
    if value is None:
        raise ValueError(f"Dynamic metadata '{name}' cannot be None.")
    arr = np.asarray(value.detach().cpu() if isinstance(value, torch.Tensor) else value, dtype=np.float64)
    if arr.ndim == 0:
        arr = arr[None]
    if arr.ndim == 2 and arr.shape[1] == 1:
        arr = arr[:, 0]
    if arr.ndim != 1:
        raise ValueError(f"Dynamic metadata '{name}' must be 0-d or 1-d, got shape {arr.shape}.")
    if arr.shape[0] != batch:
        raise ValueError(f"Dynamic metadata '{name}' length {arr.shape[0]} != batch_size {batch}.")
    return arr.tolist()


def _features(field: AbstractField, arr: np.ndarray):
    from datasets import Array2D, Array3D, Array4D, Array5D, Features, Sequence, Value

    kind = {2: Array2D, 3: Array3D, 4: Array4D, 5: Array5D}[arr.ndim]
    feat = {
        "array": kind(shape=(None, *arr.shape[1:]), dtype=np.dtype(arr.dtype).name),
        "z_sources": Sequence(Value("float64")),
        "scale_factors": Sequence(Value("float64")),
        "comoving_centers": Sequence(Value("float64")),
        "density_width": Sequence(Value("float64")),
        "mesh_size": Sequence(Value("int32"), length=3),
        "box_size": Sequence(Value("float32"), length=3),
        "observer_position": Sequence(Value("float32"), length=3),
        "halo_size": Sequence(Value("int32"), length=2),
        "status": Value("string"),
        "unit": Value("string"),
    }
    feat.update({k: Value("float32") for k in _COSMO})
    feat.update({"version": Value("int32"), "field_type": Value("string")})
    if field.nside is not None:
        feat["nside"] = Value("int32")
    if field.flatsky_npix is not None:
        feat["flatsky_npix"] = Sequence(Value("int32"), length=2)
    if field.field_size is not None:
        feat["field_size"] = Sequence(Value("float32"), length=2)
    return Features(feat)


def _row(field: AbstractField, cosmo: Cosmology, version: int):
    arr = _host_array(field)
    b = arr.shape[0]
    fs = field.field_size
    row = {
        "array": [arr],
        "z_sources": [_dynamic(field.z_sources, "z_sources", b)],
        "scale_factors": [_dynamic(field.scale_factors, "scale_factors", b)],
        "comoving_centers": [_dynamic(field.comoving_centers, "comoving_centers", b)],
        "density_width": [_dynamic(field.density_width, "density_width", b)],
        "mesh_size": [list(field.mesh_size)],
        "box_size": [list(field.box_size)],
        "observer_position": [list(field.observer_position)],
        "halo_size": [list(field.halo_size)],
        "status": [field.status.name],
        "unit": [field.unit.name],
        "field_type": [type(field).__name__],
        "version": [int(version)],
    }
    for k in _COSMO:
        row[k] = [float(getattr(cosmo, k, 0.0))]
    if field.nside is not None:
        row["nside"] = [int(field.nside)]
    if field.flatsky_npix is not None:
        row["flatsky_npix"] = [list(field.flatsky_npix)]
    if fs is not None:
        row["field_size"] = [list(fs) if isinstance(fs, (tuple, list)) else [float(fs), float(fs)]]
    return row, _features(field, arr)


def _default_device():
    return torch.device("cuda", torch.cuda.current_device()) if torch.cuda.is_available() else torch.device("cpu")


def _entry(item: dict, device) -> tuple:
    """(field, cosmology, version) of one row (catalog.py:231-338)."""
    name = item["field_type"]
    cls = _CLASSES[name]
    if cls is ParticleField:
        unit = PositionUnit[item["unit"]]
    elif cls in (SphericalKappaField, FlatKappaField):
        unit = ConvergenceUnit[item["unit"]]
    else:
        unit = DensityUnit[item["unit"]]
    arr = np.asarray(item["array"])
    single = arr.shape[0] == 1
    if single:
        arr = arr[0]

    def dyn(key):
        v = np.asarray(item[key], dtype=np.float64)
        return v[0] if single and v.ndim >= 1 else v

    def opt(key, typ):
        v = item.get(key)
        return None if v is None else tuple(typ(x) for x in v)

    fs = item.get("field_size")
    field = cls(array=torch.from_numpy(np.ascontiguousarray(arr)).to(device),
                mesh_size=tuple(int(x) for x in item["mesh_size"]),
                box_size=tuple(float(x) for x in item["box_size"]),
                observer_position=tuple(float(x) for x in item["observer_position"]),
                halo_size=tuple(int(x) for x in item["halo_size"]),
                nside=None if item.get("nside") is None else int(item["nside"]),
                flatsky_npix=opt("flatsky_npix", int),
                field_size=None if fs is None else tuple(float(x) for x in fs),
                z_sources=dyn("z_sources"), scale_factors=dyn("scale_factors"),
                comoving_centers=dyn("comoving_centers"), density_width=dyn("density_width"),
                status=FieldStatus[item["status"]], unit=unit)
    if float(item.get("Omega_nu", 0.0) or 0.0) != 0.0:
        raise ValueError("massive neutrinos (Omega_nu != 0) are not modelled by this engine")
    cosmo = Cosmology(**{k: float(item[k]) for k in _COSMO if k != "Omega_nu"})
    return field, cosmo, int(np.asarray(item["version"]).reshape(-1)[0])


class Catalog:
    """Lists of fields and their cosmologies (catalog.py:341-510). A single field with a single
    cosmology is wrapped into length-1 lists; a batched field of N simulations pairs with a list of
    N cosmologies."""

    def __init__(self, field, cosmology, version: int = CATALOG_VERSION):
        if isinstance(field, AbstractField):
            field = [field]
        if isinstance(cosmology, Cosmology):
            cosmology = [cosmology]
        self.field = list(field)
        self.cosmology = list(cosmology)
        self.version = int(version)
        if len(self.field) != len(self.cosmology):
            raise ValueError(f"field and cosmology must have the same length, got {len(self.field)} and "
                             f"{len(self.cosmology)}.")
        if not self.field:
            raise ValueError("Catalog must contain at least one field.")

    def __len__(self) -> int:
        return len(self.field)

    def __getitem__(self, key) -> "Catalog":
        if isinstance(key, int):
            return Catalog([self.field[key]], [self.cosmology[key]], self.version)
        return Catalog(self.field[key], self.cosmology[key], self.version)

    def __repr__(self) -> str:
        return (f"Catalog(n_entries={len(self.field)}, field_types={ {type(f).__name__ for f in self.field} }, "
                f"version={self.version})")

    def to_dataset(self):
        from datasets import Dataset, concatenate_datasets

        parts = []
        for f, c in zip(self.field, self.cosmology):
            row, feat = _row(f, c, self.version)
            parts.append(Dataset.from_dict(row, features=feat))
        return concatenate_datasets(parts) if len(parts) > 1 else parts[0]

    def to_parquet(self, path: str) -> None:
        self.to_dataset().to_parquet(path)

    @classmethod
    def from_dataset(cls, ds, device=None) -> "Catalog":
        if isinstance(ds, dict):
            f, c, v = _entry(ds, device or _default_device())
            return cls([f], [c], v)
        if len(ds) == 0:
            raise ValueError("Cannot reconstruct Catalog from an empty dataset.")
        device = device or _default_device()
        # the array as numpy, the metadata as Python values (numpy formatting would round the float64
        # metadata to float32)
        ds = ds.with_format("numpy", columns=["array"], output_all_columns=True)
        fields, cosmos, version = [], [], CATALOG_VERSION
        for i in range(len(ds)):
            f, c, version = _entry(ds[i], device)
            fields.append(f)
            cosmos.append(c)
        return cls(fields, cosmos, version)

    @classmethod
    def from_parquet(cls, path: str, device=None) -> "Catalog":
        from datasets import load_dataset

        return cls.from_dataset(load_dataset("parquet", data_files=path, split="train"), device=device)


def catalog_from_results(results, requests) -> Catalog:
    """Catalog of served forward simulations: one DensityField (the final painted particle counts)
    per result, with the request's cosmology (serving.SimResult / SimRequest)."""
    by_id = {r.rid: r for r in requests}
    fields, cosmos = [], []
    for res in results:
        req = by_id[res.rid]
        n = int(res.mesh)
        box = req.box_size()
        fields.append(DensityField(array=res.density, mesh_size=(n, n, n), box_size=(box,) * 3,
                                   status=FieldStatus.DENSITY_FIELD, unit=DensityUnit.DENSITY,
                                   z_sources=1.0 / req.a1 - 1.0, scale_factors=req.a1, comoving_centers=0.0,
                                   density_width=0.0))
        cosmos.append(req.cosmology())
    return Catalog(fields, cosmos)
