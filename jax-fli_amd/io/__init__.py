"""On-disk catalogs of simulation outputs (src/jax_fli/io/catalog.py)."""

from .catalog import CATALOG_VERSION, Catalog, catalog_from_results

__all__ = ["Catalog", "CATALOG_VERSION", "catalog_from_results"]
