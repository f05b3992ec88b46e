"""Lenia individuals of a quality-diversity search (leniax/lenia.py): the genotype decoding that
turns a solution's raw parameters into the configuration the step runs."""
import copy
from typing import Dict, List, Tuple

from . import utils as lx_utils


class LeniaIndividual(object):
    """leniax/lenia.py:8-77: configuration, random key and raw genotype values in [0, 1]."""

    def __init__(self, config: Dict, rng_key, params: List = ()):
        self.qd_config = copy.deepcopy(config)
        self.rng_key = rng_key
        self.params = list(params)
        self.fitness = 0.
        self.features: List[float] = []
        if 'genotype' in self.qd_config:
            for gene in self.get_genotype():
                lx_utils.get_param(self.qd_config, gene['key'])  # genes may only update existing values

    def set_init_props(self, rng_key, best_init_idxs: List[int]):
        key = rng_key.tolist() if hasattr(rng_key, 'tolist') else rng_key
        self.qd_config['algo']['init_rng_key'] = key
        self.qd_config['algo']['best_init_idxs'] = best_init_idxs

    def set_cells(self, cells: str):
        self.qd_config['run_params']['cells'] = cells

    def set_init_cells(self, init_cells: str):
        self.qd_config['run_params']['init_cells'] = init_cells

    def get_config(self) -> Dict:
        if 'genotype' in self.qd_config:
            genotype = self.get_genotype()
            raw_values = [round(v, 8) for v in self.params]  # leniax/lenia.py:68-69
            assert len(raw_values) == len(genotype)
            return update_config(self.qd_config, get_update_config(genotype, raw_values))
        return self.qd_config

    def get_genotype(self):
        return self.qd_config['genotype']


def update_config(config: Dict, to_update: Dict) -> Dict:
    """leniax/lenia.py:80-98."""
    new_config = copy.deepcopy(config)
    if 'kernels_params' in to_update:
        for i, kernel in enumerate(to_update['kernels_params']):
            new_config['kernels_params'][i].update(kernel)
    if 'world_params' in to_update:
        new_config['world_params'].update(to_update['world_params'])
    return new_config


def get_update_config(genotype: List[Dict], raw_values: List) -> Dict:
    """leniax/lenia.py:101-128: each gene's raw value scaled into its domain (float, int, choice)."""
    to_update: Dict = {}
    for gene, raw_val in zip(genotype, raw_values):
        domain, val_type = gene['domain'], gene['type']
        if val_type == 'float':
            val = float(linear_scale(raw_val, domain))
        elif val_type == 'int':
            val = int(linear_scale(raw_val, (domain[0], domain[1] + 1)) - 0.5)
        elif val_type == 'choice':
            val = domain[int(linear_scale(raw_val, (0, len(domain))) - 0.5)]
        else:
            raise ValueError(f'type {val_type} unknown')
        lx_utils.set_param(to_update, gene['key'], val)
    return to_update


def linear_scale(raw_value: float, domain: Tuple[float, float]) -> float:
    """leniax/lenia.py:131-143."""
    return domain[0] + (domain[1] - domain[0]) * raw_value


def log_scale(raw_value: float, domain: Tuple[float, float]) -> float:
    """leniax/lenia.py:146-158."""
    return domain[0] * (domain[1] / domain[0]) ** raw_value
