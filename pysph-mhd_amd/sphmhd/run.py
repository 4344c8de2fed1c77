"""Command line (reference: src/run.py:16-43): ``python -m sphmhd.run CONFIG``
from the pysph-mhd_amd directory, with the reference's TOML schema."""
from __future__ import annotations

import argparse
from pathlib import Path

from .cases.apps import AlfvenWave2D, MHDBlast3D, MHDRotor2D, OrszagTang2D
from .cases.reference import CASES as REFERENCE_CASES
from .config import load_config, pysph_arguments

# the reference's twelve case names (its TOMLs run unchanged) and the
# benchmark workloads' 2-D and cubic-lattice variants
CASES = {
    **REFERENCE_CASES,
    "orstang2d": OrszagTang2D,
    "mhdrotor2d": MHDRotor2D,
    "mhdblast3d": MHDBlast3D,
    "alfven2d": AlfvenWave2D,
}


def main(argv=None):
    parser = argparse.ArgumentParser()
    parser.add_argument("config", type=Path)
    parser.add_argument("--max-steps", type=int, default=None)
    arguments = parser.parse_args(argv)
    config_types = {name: application_type.config_type for name, application_type in CASES.items()}
    config = load_config(arguments.config, config_types)
    application = CASES[config.case.name](config)
    extra = ["--max-steps", str(arguments.max_steps)] if arguments.max_steps is not None else []
    application.run(pysph_arguments(config.execution) + extra)
    return application


if __name__ == "__main__":
    main()
