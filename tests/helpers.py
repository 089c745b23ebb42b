"""Test helpers: convert product-side media and ray batches to the oracle's
layout (test infrastructure only)."""
from __future__ import annotations

import numpy as np

from oracle import pyoracle as po


def oracle_desc(med):
    d = po.make_desc(
        cov_kind=med.cov_kind, sigma=med.sigma, length_scale=med.length_scale, kernel_scale=med.local_scale,
        aniso=med.aniso, ns_multires=int(med.ns_multires), ns_ls_field=med.ns_ls_field,
        ns_ls_params=med.ns_ls_params, ns_ls_max=med.ns_ls_max, ns_var=med.ns_var, mean_kind=med.mean_kind,
        mean_radius=med.mean_radius, mean_center=med.mean_center, mean_half=med.mean_half,
        plane_normal=med.plane_normal, plane_offset=med.plane_offset, global_seed=med.seed,
        impulse_density=med.impulse_density, single_realization=int(med.single_realization),
        ctxt=med.correlation_context, one_d=int(med.one_d), scheme=med.scheme,
        correlation_xy=int(med.correlation_xy), step_size=med.step_size, min_step=med.min_step,
        max_bounce=med.max_bounces, grid=med.mean_grid, grid_offset=med.mean_offset, grid_scale=med.mean_scale,
        grid_volume=med.mean_volume, iso3d=int(med.isotropic_3D_sampling and not med.one_d),
        sdf_function=med.sdf_function, sdf_inverse_transform=med.sdf_inverse_transform, sdf_min=med.sdf_min,
        sdf_scale=med.sdf_scale, sdf_offset=med.sdf_offset, ns_var_field=med.ns_var_field,
        ns_var_noise=med.ns_var_noise, ns_var_noise_params=med.ns_var_noise_params, ns_ls_noise=med.ns_ls_noise,
        ns_ls_noise_params=med.ns_ls_noise_params,
        aniso_mtx=np.asarray(med.aniso_mtx, np.float32).reshape(3, 3) if med.use_aniso_mtx else None)
    return d


def oracle_field(med):
    return po.Field(oracle_desc(med))


def oracle_rays(rays, med, scene_seed=None):
    n = rays["ox"].shape[0]
    r = np.zeros(n, po.RAY_DTYPE)
    r["origin"] = np.stack([rays["ox"], rays["oy"], rays["oz"]], 1)
    r["dir"] = np.stack([rays["dx"], rays["dy"], rays["dz"]], 1)
    r["near_t"], r["far_t"], r["jitter"] = rays["near_t"], rays["far_t"], rays["jitter"]
    r["pss"] = rays["pss"]
    r["scene_seed"] = rays.get("scene_seed", 0) if scene_seed is None else scene_seed
    r["first_scatter"] = (rays["flags"] & 1).astype(np.int32)
    r["last_val"] = rays["prev_value"]
    r["last_aniso"] = np.stack([rays["prev_gx"], rays["prev_gy"], rays["prev_gz"]], 1).astype(np.float64)
    r["bounce"] = 0
    return r
