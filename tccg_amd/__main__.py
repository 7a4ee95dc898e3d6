"""python -m tccg_amd ...  (the tccg command line, see tccg_amd.cli)."""
import sys

from .cli import main

sys.exit(main())
