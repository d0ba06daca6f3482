"""vxp — MI355X step pipeline for zig-voxel-physics (Python host mirror).

The compute path is libvxp.so (HIP kernels for gfx950 behind the C ABI in
include/vxp.h).  This package binds it with ctypes, mirrors the reference's
VoxelPhysicsSystem API and generates the seeded benchmark scenes.
"""
from ._abi import (Config, Handle, Lib, VxpError, default_config, oracle_lib, product_lib,
                   STATIC, KINEMATIC, DYNAMIC, EMPTY, CORNER, EDGE, FACE, INSIDE,
                   SOLVE_ORDER_COLOR, SOLVE_ORDER_CANONICAL, voxel_byte)
from .system import VoxelCollider, VoxelPhysicsSystem, from_scene

__all__ = ["Config", "Handle", "Lib", "VxpError", "default_config", "oracle_lib", "product_lib",
           "VoxelCollider", "VoxelPhysicsSystem", "from_scene", "STATIC", "KINEMATIC", "DYNAMIC",
           "EMPTY", "CORNER", "EDGE", "FACE", "INSIDE", "SOLVE_ORDER_COLOR",
           "SOLVE_ORDER_CANONICAL", "voxel_byte"]
