"""The build recipe loads in a checkout where nothing is built yet.

The package magpy_amd imports its compiled binding on import, so the build
script must be reachable without importing the package: __graft_entry__
loads magpy_amd/build.py by path.  This copies the sources (no built
libraries) into an empty directory and loads the recipe there in a fresh
interpreter; the compile itself is exercised by every build.
"""
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_build_recipe_loads_without_the_binding(tmp_path):
    shutil.copy(os.path.join(ROOT, "__graft_entry__.py"), tmp_path)
    shutil.copytree(os.path.join(ROOT, "magpy_amd"), tmp_path / "magpy_amd",
                    ignore=shutil.ignore_patterns("*.so", "*.so.tmp*", "__pycache__"))
    assert not any(p.suffix == ".so" for p in (tmp_path / "magpy_amd").iterdir())
    code = (
        "import os, sys\n"
        "import __graft_entry__ as g\n"
        "B = g._load_by_path('b', os.path.join('magpy_amd', 'build.py'))\n"
        "assert 'magpy_amd' not in sys.modules\n"
        "assert B.LIB == os.path.abspath(os.path.join('magpy_amd', 'libmagpy_amd.so'))\n"
        "assert all(os.path.exists(os.path.join('magpy_amd', 'csrc', s)) for s in B.SOURCES)\n"
        "try:\n"
        "    import magpy_amd  # noqa: F401\n"
        "except ImportError:\n"
        "    print('package import fails loudly without its binding')\n"
        "else:\n"
        "    raise SystemExit('package imported without its binding')\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "fails loudly" in r.stdout
