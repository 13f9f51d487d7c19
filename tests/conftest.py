import os
import sys

# Calls on several streams (per-frame calls, the serving loop's batches)
# overlap on the device only when each of the process's streams has a
# hardware queue of its own; the runtime's pool holds GPU_MAX_HW_QUEUES of
# them (4 by default), read when the runtime starts.
os.environ["GPU_MAX_HW_QUEUES"] = "8"

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (HIP device)")
