"""Synthetic volumes (API of the reference's ``pyvr.datasets`` plus the
seeded blob-field family used by the large render workloads)."""

from .synthetic import blob_field, blob_params, compute_normal_volume, create_sample_volume

__all__ = ["create_sample_volume", "compute_normal_volume", "blob_field", "blob_params"]
