"""The library and torch share one HIP runtime whatever is used first: a
process that calls the library before touching torch's GPU must still see
the device from torch (torch bundles its own libamdhip64)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.gpu
def test_library_first_then_torch():
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "from ducc_amd import _lib\n"
        "assert _lib.lib().dmx_device_count() >= 1\n"
        "import torch\n"
        "assert torch.cuda.is_available()\n"
        "x = torch.ones(4, device='cuda')\n"
        "assert float(x.sum()) == 4.0\n"
        "print('ok')\n" % ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr
