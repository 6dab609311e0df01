"""Polygon geometry: exact-coverage rasterizer and painter's-order scene
compositor feeding the engine's eps / pole / chi3 maps."""
from .polygons import Polygon
from .raster import coverage, coverage_array, normals
from .scene import RasterizedMaterials, kottke_pair, rasterize_scene

__all__ = ["Polygon", "coverage", "coverage_array", "normals", "RasterizedMaterials", "kottke_pair",
           "rasterize_scene"]
