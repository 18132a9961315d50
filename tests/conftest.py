import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT / "squares-in-disk_amd", ROOT / "oracle", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


# Published totals (num_cuboids, num_leaf_cuboids) of the reference prover,
# prover_output_23_11_2020.txt:2-34 (also README.md:46-69).
PUBLISHED = {
    "top1": (128, 128),
    "top2": (168, 168),
    "one": (2869888, 2858678),
    "two": (7181696, 7069587),
    "three": (3549856, 3440786),
    "four": (37245328, 35026303),
    "five": (3392, 3188),
    "six": (14848, 13928),
    "seven": (1808, 1703),
    "more1": (1056154896, 1022752013),
    "more2": (123127968, 115680431),
}


@pytest.fixture(scope="session")
def program_text():
    import sqd
    cache = {}

    def get(key):
        if key not in cache:
            cache[key] = sqd.Program.from_lemma(key).text()
        return cache[key]
    return get
