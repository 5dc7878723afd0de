"""The configuration fields the interferogram path reads, with the
reference's names (_config.py:15-152; defaults config/default.yaml:22-47).

``load_config`` reads the reference's YAML layout with a safe loader and
keeps only the sections used here (io, proc, filter); unknown keys are
ignored rather than rejected, so a full s1proc config file loads.
"""

from dataclasses import dataclass, field, fields
from pathlib import Path
from typing import Optional

import yaml


@dataclass
class IoConfig:
    slc_path: str = "slc"
    amp_path: str = "amp"
    ifg_path: str = "igrams"
    ifg_corr_path: str = "igrams_corr"
    dem_file: str = ""
    rsc_file: str = ""
    img_pair_file: str = "ifglist"
    multilook_rsc_file: str = ""


@dataclass
class ProcessingConfig:
    rowlook: int = 10
    collook: int = 20
    wavelength: float = 0.0554658
    io_workers: Optional[int] = None
    cpu_workers: Optional[int] = None
    gpu_workers: Optional[int] = None
    max_slots: Optional[int] = None
    streams_per_gpu: Optional[int] = None


@dataclass
class FilteringParams:
    window_size: int = 32
    goldstein_alpha: float = 0.5


@dataclass
class FilteringConfig:
    enable: bool = False
    method: str = "goldstein"
    parameters: FilteringParams = field(default_factory=FilteringParams)


@dataclass
class S1Config:
    io: IoConfig = field(default_factory=IoConfig)
    proc: ProcessingConfig = field(default_factory=ProcessingConfig)
    filter: FilteringConfig = field(default_factory=FilteringConfig)


def _fill(cls, data):
    data = data or {}
    kw = {}
    for f in fields(cls):
        if f.name not in data:
            continue
        v = data[f.name]
        if f.name == "parameters" and cls is FilteringConfig:
            v = _fill(FilteringParams, v)
        kw[f.name] = v
    return cls(**kw)


def load_config(config_file, relative_path: bool = True) -> S1Config:
    """Load the io / proc / filter sections of an s1proc YAML config
    (_config.py:337-374); io paths are made relative to the file."""
    with open(config_file, "r") as f:
        raw = yaml.safe_load(f) or {}
    cfg = S1Config(io=_fill(IoConfig, raw.get("io")), proc=_fill(ProcessingConfig, raw.get("proc")),
                   filter=_fill(FilteringConfig, raw.get("filter")))
    if relative_path:
        root = Path(config_file).parent
        for f in fields(IoConfig):
            v = getattr(cfg.io, f.name)
            if v is not None and v != "":
                setattr(cfg.io, f.name, str(root / v))
    return cfg
