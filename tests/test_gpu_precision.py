"""FP32-shifted populations against FP64 on the D3Q19 MRT 256^3 droplet
(sphere R = 40, 5000 steps, the benchmark's physics): the density profile
(liquid and gas densities, largest pointwise difference), the surface
tension from Laplace's law with the equimolar radius, the spurious-current
magnitude (max Mach) and the mass drift of each precision, within stated
bounds. The FP64 run is the path pinned to the oracle."""
import pytest

pytestmark = pytest.mark.gpu


def test_fp32_shifted_vs_fp64_droplet_5000_steps():
    import json
    import bench
    r = bench.run_precision_study(5000)
    print(json.dumps(r), flush=True)
    d = r['fp32s_vs_fp64']
    assert d['rel_n_l'] <= 2e-5          # liquid density
    assert d['rel_n_g'] <= 1e-4          # gas density
    assert d['max_abs_dn'] <= 0.05       # largest difference, at the interface
    assert d['rel_sigma'] <= 2e-3        # surface tension (Laplace, equimolar radius)
    assert d['rel_max_mach'] <= 1e-2     # spurious currents
    assert r['fp64']['mass_rel_drift'] <= 1e-12
    assert r['fp32s']['mass_rel_drift'] <= 1e-4
