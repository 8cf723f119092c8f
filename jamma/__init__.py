"""Import alias: ``jamma`` resolves to this repository's ``jamma_amd`` package.

A caller of the reference writes ``from jamma.kinship import compute_centered_kinship``,
``from jamma.lmm import run_lmm_association_streaming``, ``from jamma.lmm.runner_jax
import run_lmm_association_jax`` or ``import jamma; jamma.gwas(...)``: with this
directory on the import path those imports load the device implementation unchanged.
Every ``jamma.<x>`` module is the ``jamma_amd.<x>`` module of the same name (the same
object in ``sys.modules``); ``jamma.lmm.runner_jax`` is the module that holds the
in-memory runner here. Reference modules outside the association hot path (the CLI,
``jamma.core``'s JAX / threading / memory configuration, validation tooling) have no
counterpart and raise ImportError.
"""

from __future__ import annotations

import importlib
import importlib.abc
import importlib.util
import sys

import jamma_amd

# reference module -> module of this package where it differs by name
_RENAMED = {
    "jamma.lmm.runner_jax": "jamma_amd.lmm.runner_streaming",
}


class _AliasLoader(importlib.abc.Loader):
    def __init__(self, target: str):
        self.target = target

    def create_module(self, spec):
        return importlib.import_module(self.target)

    def exec_module(self, module):  # already executed as the target module
        return None


class _AliasFinder(importlib.abc.MetaPathFinder):
    def find_spec(self, fullname, path=None, target=None):
        if not fullname.startswith("jamma."):
            return None
        real = _RENAMED.get(fullname, "jamma_amd." + fullname[len("jamma."):])
        try:
            if importlib.util.find_spec(real) is None:
                return None
        except ModuleNotFoundError:
            return None
        is_pkg = importlib.util.find_spec(real).submodule_search_locations is not None
        return importlib.util.spec_from_loader(fullname, _AliasLoader(real), is_package=is_pkg)


if not any(isinstance(f, _AliasFinder) for f in sys.meta_path):
    sys.meta_path.insert(0, _AliasFinder())

from jamma_amd import *  # noqa: E402,F401,F403
from jamma_amd import GWASResult, __version__, gwas  # noqa: E402,F401

__all__ = list(jamma_amd.__all__)
