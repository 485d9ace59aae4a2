"""Host-side checks of the 3D interface that need no GPU: configurations
vv3_create rejects before it touches a device, and the field layout."""
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "codesthesis_dir_amd", "libvvmi.so")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libvvmi.so not built")


@pytest.mark.parametrize("bad", [dict(np=15), dict(nr=3), dict(nz=2), dict(dt=0.0), dict(rmax=0.4),
                                 dict(Re=0.0), dict(np=2)])
def test_create_rejects_bad_configuration(bad):
    from codesthesis_dir_amd._lib import VV_EINVAL, VVError
    from codesthesis_dir_amd.velvor3 import Annulus3D, Annulus3DConfig
    p = dict(nr=8, nz=8, np=8, dt=1e-4, Ra=1e4, Re=30.0, Ptl=1.0, Gr=5.0)
    p.update(bad)
    with pytest.raises(VVError) as e:
        Annulus3D(Annulus3DConfig(**p))
    assert e.value.code == VV_EINVAL


def test_field_names_follow_the_header_enum():
    import re
    from codesthesis_dir_amd.velvor3 import FIELDS
    src = open(os.path.join(ROOT, "include", "vvmi3.h")).read()
    enum = re.search(r"enum vv3_field \{(.*?)\};", src, re.S).group(1)
    names = [t.strip() for t in enum.replace("\n", " ").split(",") if t.strip()]
    names = [n.split("=")[0].strip() for n in names if n.strip() != "VV3_NFIELD"]
    assert names == ["VV3_" + f.upper() for f in FIELDS]
