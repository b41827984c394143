import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def mesh_dir(tmp_path_factory):
    """Small O-grid pipe: 4x4 core, 2 rings, 6 slabs (1,728 tets)."""
    from nodals_amd import meshgen

    d = tmp_path_factory.mktemp("meshes")
    return {
        "small": str(meshgen.write_pipe_case(d / "small", core=4, ring=2, nslab=6)),
        "medium": str(meshgen.write_pipe_case(d / "medium", core=6, ring=3, nslab=12)),
        "helix": str(meshgen.write_helix_case(d / "helix", core=4, ring=2, nslab=48)),
    }


def pytest_collection_modifyitems(config, items):
    pass
