"""Config loading with the reference's merge order (train_per_scene.py:243-247):
defaults.yaml <- scene yaml <- CLI dotlist (``a.b=v``), without OmegaConf.

Values are parsed with yaml.safe_load, so ``[1,2]`` is a list and ``1e-3`` a
float, as OmegaConf would give them.
"""

import copy
from collections.abc import Mapping, Sequence

import yaml


class AttrDict(dict):
    """dict with attribute access, nested (stands in for an OmegaConf DictConfig)."""

    def __getattr__(self, k):
        try:
            return self[k]
        except KeyError as e:
            raise AttributeError(k) from e

    def __setattr__(self, k, v):
        self[k] = v

    def __deepcopy__(self, memo):
        return AttrDict({k: copy.deepcopy(v, memo) for k, v in self.items()})

    @staticmethod
    def wrap(d):
        """Nested AttrDicts from any mapping (a dict, an AttrDict, or a
        mapping-like config such as an OmegaConf DictConfig read from a
        reference checkpoint) and lists/tuples of them."""
        if isinstance(d, Mapping):
            return AttrDict({k: AttrDict.wrap(v) for k, v in d.items()})
        if isinstance(d, (list, tuple)) or _is_listconfig(d):
            return [AttrDict.wrap(v) for v in d]
        return d

    def to_dict(self):
        return plain_cfg(self)


def _is_listconfig(d):
    # OmegaConf's ListConfig is a Sequence but neither a list nor a str
    return isinstance(d, Sequence) and not isinstance(d, (str, bytes, list, tuple))


def plain_cfg(d):
    """A config as plain nested dicts and lists (what a checkpoint stores):
    accepts AttrDicts, dicts and mapping-like configs."""
    if isinstance(d, Mapping):
        return {k: plain_cfg(v) for k, v in d.items()}
    if isinstance(d, (list, tuple)) or _is_listconfig(d):
        return [plain_cfg(v) for v in d]
    return d


def _parse_scalar(s):
    # YAML 1.1 (PyYAML) does not read '1e-3' as a float; OmegaConf does.
    v = yaml.safe_load(s)
    if isinstance(v, str):
        try:
            return float(v)
        except ValueError:
            return v
    return v


def _fix_floats(d):
    if isinstance(d, dict):
        return {k: _fix_floats(v) for k, v in d.items()}
    if isinstance(d, list):
        return [_fix_floats(v) for v in d]
    if isinstance(d, str):
        try:
            if any(c in d for c in "eE.") and not d.strip().startswith(("'", '"')):
                return float(d)
        except ValueError:
            pass
    return d


def merge(base, other):
    """Recursive merge: values of `other` override `base`."""
    out = AttrDict.wrap(copy.deepcopy(base))
    for k, v in (other or {}).items():
        if isinstance(v, dict) and isinstance(out.get(k), dict):
            out[k] = merge(out[k], v)
        else:
            out[k] = AttrDict.wrap(copy.deepcopy(v))
    return out


def from_dotlist(items):
    out = {}
    for it in items:
        if "=" not in it:
            raise ValueError(f"expected key=value, got {it!r}")
        key, val = it.split("=", 1)
        node = out
        parts = key.strip().split(".")
        for p in parts[:-1]:
            node = node.setdefault(p, {})
        node[parts[-1]] = _parse_scalar(val)
    return AttrDict.wrap(out)


def load_yaml(path):
    with open(path) as f:
        return AttrDict.wrap(_fix_floats(yaml.safe_load(f) or {}))


def load_config(*paths, dotlist=()):
    cfg = AttrDict()
    for p in paths:
        cfg = merge(cfg, load_yaml(p) if isinstance(p, str) else AttrDict.wrap(p))
    if dotlist:
        cfg = merge(cfg, from_dotlist(dotlist))
    return cfg
