"""pip packaging of parament_amd.

`pip install --no-build-isolation .` compiles libparament.so with hipcc (for gfx950, or the
architecture in PARAMENT_OFFLOAD_ARCH) and installs it inside the package, where
parament_amd.paramentlib loads it; the C header goes to <prefix>/include.  The build recipe is
parament_amd/build.py, loaded here by path so that the package itself (which loads the library
at import) is not imported before the library exists.
"""
import importlib.util
import os

from setuptools import setup
from setuptools.command.build_py import build_py

ROOT = os.path.dirname(os.path.abspath(__file__))


class BuildWithHipcc(build_py):
    """Compile libparament.so into the source package before the files are collected."""

    def run(self):
        spec = importlib.util.spec_from_file_location("parament_amd_build", os.path.join(ROOT, "parament_amd", "build.py"))
        recipe = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(recipe)
        recipe.build(verbose=True)
        super().run()


setup(
    name="parament_amd",
    version="0.1.0",
    description="Piecewise-constant Hamiltonian propagators (Chebyshev equiprop) on MI355X: C ABI, Python class, "
                "mixed precision, adjoint gradients, ensembles",
    packages=["parament_amd", "parament_amd.test"],
    package_data={"parament_amd": ["libparament.so"]},
    data_files=[("include", ["include/parament.h"])],
    python_requires=">=3.9",
    install_requires=["numpy"],
    cmdclass={"build_py": BuildWithHipcc},
    zip_safe=False,
)
