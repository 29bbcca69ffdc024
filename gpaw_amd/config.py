"""Run-time switches mirroring gpaw/__init__.py:52-53,186-216.

debug             extra argument checks in the operator wrappers
                  (the reference's --debug, gpaw/fd_operators.py:97-121)
extra_parameters  the reference's --gpaw=... dictionary; 'newgucstencil'
                  selects NewGUCLaplace (default) or GUCLaplace
"""
import os

debug = os.environ.get('GPAW_AMD_DEBUG', '0') not in ('', '0')
extra_parameters = {}
