"""Ordered YAML input (yaml_parser.py:7-50 of the reference).

Element order in a lattice file is significant (its first key is
``step_size``; the elements follow in beam order), so mappings load into
OrderedDict.  A string names a file (relative to the working directory);
anything else is read as a stream.  A parse failure prints the error and
exits with status 1, as the reference does.
"""
from __future__ import annotations

import os
import sys
from collections import OrderedDict

import yaml


def full_path(path):
    return os.path.abspath(os.path.expandvars(path))


def ordered_load(stream):
    class _Loader(yaml.SafeLoader):
        pass

    def _mapping(loader, node):
        loader.flatten_mapping(node)
        return OrderedDict(loader.construct_pairs(node))

    _Loader.add_constructor(yaml.resolver.BaseResolver.DEFAULT_MAPPING_TAG, _mapping)
    return yaml.load(stream, _Loader)


def parse_yaml(source, base_dir=None):
    if isinstance(source, str):
        cands = [full_path(source)]
        if base_dir is not None:
            cands.append(os.path.join(base_dir, source))
        for c in cands:
            if os.path.exists(c):
                with open(c) as f:
                    return ordered_load(f)
        raise FileNotFoundError(f"YAML input not found: {source}")
    try:
        d = ordered_load(source)
        assert isinstance(d, dict), f"ERROR: parsing unsuccessful, could not read {source}"
        return d
    except Exception as ex:   # reference: print and sys.exit(1)
        print(ex)
        sys.exit(1)
