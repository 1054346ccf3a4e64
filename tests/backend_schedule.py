"""RadiationModel::runBand's host pass loop driven through the call-by-call backend entry points.

This is the caller-side sequence of plugins/radiation/src/RadiationModel.cpp, restated in Python
over ``native.Backend`` so that a GPU test can check the call-by-call interface against the
device-resident ``rt_run_band`` on the same seeds:

- zeroRadiationBuffers, uploadSourceFluxes, launchDirectRays (RadiationModel.cpp:3804-3976);
- scattered direct energy to radiation_out, one-sided primitives combined (:3999-4036);
- emission added on the host, primary diffuse launch on the top then the bottom face (:4040-4209);
- copyScatterToRadiation when nothing ran directly (:4225-4234);
- scattering iterations with per-band retirement (:4236-4356);
- radiation_flux = radiation_in + scatter_top + scatter_bottom (:4631-4654).
"""
from __future__ import annotations

import numpy as np

SIGMA = 5.6703744e-8


def run_band_via_backend(be, req, twosided, band_map=None, nbands_global=None):
    """Runs one runBand through ``be`` (a native.Backend with geometry, BVH, materials and sources
    already set) for the bands of ``req`` (a native.BandRequest).  Returns radiation_flux [P][B]."""
    P, B = req.P, req.B
    Bg = B if nbands_global is None else nbands_global
    S = req.sdir.shape[0]
    depth = [int(x) for x in req.depth]
    emission_b = [bool(x) for x in req.emission]
    dflux = req.dflux.astype(np.float32)

    be.zero_radiation_buffers(B)
    rundirect = bool(S > 0 and np.any(req.sflux > 0.0))
    if S > 0:
        be.upload_source_fluxes(req.sflux)
    if S > 0 and rundirect:
        be.launch_direct_rays(req.direct_rays, req.direct_seed, Bg, band_map=band_map)

    scattering = any(d > 0 for d in depth)
    emission = any(emission_b)
    diffuse = bool(np.any(dflux > 0.0))
    flux_top = np.zeros((P, B), np.float32)
    flux_bottom = np.zeros((P, B), np.float32)
    if scattering and rundirect:
        r = be.get_radiation_results()
        flux_top = r["scatter_buff_top"].copy()
        flux_bottom = r["scatter_buff_bottom"].copy()
        one = np.asarray(twosided) == 0
        tot = flux_top[one] + flux_bottom[one]
        flux_top[one] = tot
        flux_bottom[one] = tot
        be.upload_radiation_out(flux_top, flux_bottom)
        be.zero_scatter_buffers()

    if emission or diffuse:
        if emission:
            eps = np.ones((P, B), np.float32) if req.eps is None else req.eps
            temp = np.full(P, 300.0, np.float32) if req.temp is None else req.temp
            two = np.asarray(twosided) != 0
            for b in range(B):
                if not emission_b[b]:
                    continue
                e = eps[:, b].astype(np.float64)
                if depth[b] == 0:
                    e = np.where(e != 1.0, 1.0, e)
                T = np.where(temp < 0, 300.0, temp).astype(np.float64)
                out = (SIGMA * e * T ** 4).astype(np.float32)
                flux_top[:, b] += out
                flux_bottom[two, b] += out[two]
        for face in (1, 0):
            be.launch_diffuse_rays(req.diffuse_rays, int(req.dseeds[0]), Bg, face, diffuse_flux=dflux,
                                   band_map=band_map, radiation_out=(flux_top, flux_bottom),
                                   sky_extinction=getattr(req, "sky_ext", None),
                                   sky_peak_dir=getattr(req, "sky_peak", None), sky_norm=getattr(req, "sky_norm", None))

    if scattering and (emission or diffuse) and not rundirect:
        be.copy_scatter_to_radiation()
        be.zero_scatter_buffers()

    if scattering and (emission or diffuse or rundirect):
        zero = np.zeros(B, np.float32)
        active = np.ones(B, np.uint8)
        for s in range(max(depth)):
            for b in range(B):
                if s + 1 > depth[b]:
                    active[b] = 0
            if s > 0 or (emission and rundirect):
                be.copy_scatter_to_radiation()
            be.zero_scatter_buffers()
            r = be.get_radiation_results()
            out = (r["radiation_out_top"], r["radiation_out_bottom"])
            for face in (1, 0):
                be.launch_diffuse_rays(req.diffuse_rays, int(req.dseeds[1 + s]), Bg, face, diffuse_flux=zero,
                                       band_map=band_map, band_active=active, radiation_out=out)

    r = be.get_radiation_results()
    return (r["radiation_in"] + r["scatter_buff_top"]) + r["scatter_buff_bottom"]
