import os
import sys

# torch first: its bundled HIP runtime (soname libamdhip64.so.7) is then the one
# the libraries under test bind to. Loaded after them, torch brings in its own
# copy beside the system one, and two HIP runtimes in one process can leave the
# second without a visible GPU.
import torch  # noqa: F401,E402

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(REPO, "tests"), os.path.join(REPO, "sbwt-search_amd", "bindings")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(REPO, "tests", "golden")
FIXTURES = os.path.join(GOLDEN, "reference_fixtures")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels on an MI355X (gfx950)")
