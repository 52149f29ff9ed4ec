"""Host-side checks of the torch.autograd binding (no GPU): parameter layout conversion."""
import numpy as np

from falcor_amd import autograd as A
from falcor_amd import scene as S


def test_material_param_layout_round_trip():
    """Records <-> the reference's 20-float StandardMaterialParamLayout (baseColor 0-2, metallic 3,
    roughness 4, ior 5, transmissionColor 6-8, diffuseTransmission 9, specularTransmission 10,
    emissive 11-13, emissiveFactor 14)."""
    r = np.stack([S.standard_material(base=(0.2, 0.3, 0.4), roughness=0.5, metallic=0.1, ior=1.4, emissive=(1, 2, 3),
                                      emissive_factor=2.0, double_sided=True, texture=3),
                  S.pbrt_diffuse_material((0.5, 0.6, 0.7)),
                  S.standard_material(specular_transmission=0.75, diffuse_transmission=0.25, transmission=(0.9, 0.8, 0.7))])
    p = A.material_params(r)
    assert p.shape == (3, A.PARAM_COUNT)
    np.testing.assert_array_equal(p[0, [0, 1, 2, 3, 4, 5, 11, 12, 13, 14]],
                                  np.float32([0.2, 0.3, 0.4, 0.1, 0.5, 1.4, 1, 2, 3, 2]))
    np.testing.assert_array_equal(p[0, 6:11], [1, 1, 1, 0, 0])  # default transmission colour 1
    np.testing.assert_array_equal(p[2, 6:11], np.float32([0.9, 0.8, 0.7, 0.25, 0.75]))
    np.testing.assert_array_equal(A.material_records(p, r), r)
    q = p.copy()
    q[1, 0:3] = (0.1, 0.2, 0.3)
    r2 = A.material_records(q, r)
    np.testing.assert_array_equal(r2[1, S.M_BASE:S.M_BASE + 3], np.float32([0.1, 0.2, 0.3]))
    np.testing.assert_array_equal(r2[1, S.M_TYPE], r[1, S.M_TYPE])
