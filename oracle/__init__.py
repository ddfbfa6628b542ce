"""CPU oracle for the Factor Fields field-evaluation and volume-rendering path.

TEST INFRASTRUCTURE ONLY.  This package is a CPU restatement of the reference
algorithm (autonomousvision/factor-fields, ``models/FactorFields.py``,
``models/sh.py``, ``dataLoader/ray_utils.py``, ``utils.py``) written with the
same ATen operations and the same operation order as the reference, so that it
reproduces the reference's FP32 semantics on the same inputs.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only to check results or as the timed CPU baseline -- never as part of
the product path, which must fail loudly when its HIP library is missing.

How each function is pinned (checked against something other than itself):

* ``grid_mapping``      -- closed forms, periodicity / range invariants and the
                           identity special case (tests/test_oracle_field.py).
* ``dct_dict``          -- closed-form DCT rows, selected-row tables and unit
                           row norms (tests/test_oracle_field.py).
* ``setup_params``      -- the parameter counts the reference notebooks printed
                           (1,429,376 for configs/image.yaml; 5,342,274 for
                           configs/sdf.yaml) and hand arithmetic
                           (tests/test_oracle_field.py).
* ``grid_sample`` uses  -- an independent numpy brute-force bi/trilinear and
                           nearest interpolator (tests/npref.py), partition of
                           unity, FP64 central differences for the gradients
                           (tests/test_oracle_field.py).
* VM coefficient        -- new semantics (the reference's ``vm`` coefficient
                           cannot be combined with a grid basis); pinned by the
                           same numpy brute force on planes and lines.
* ``MLPMixer``          -- FP64 matrix products in numpy (tests/npref.py).
* sampler / occupancy   -- closed forms with injected jitter; exact integer
                           properties (monotone t, in-box masks); brute-force
                           trilinear occupancy (tests/test_oracle_render.py).
* ``raw2alpha``/composite -- telescoping identity sum(w) + T_last = prod
                           (with the +1e-10 term), single-sample and sigma = 0
                           special cases, FP64 central differences
                           (tests/test_oracle_render.py).
* ``eval_sh``           -- orthonormality of the degree-2 real SH basis by
                           quadrature on the sphere.
* ``sample_pdf``        -- uniform-weights special case (stratified lerp) and a
                           pure-Python inverse-CDF loop on tiny inputs.
* proposal              -- explicit-overlap loops for the bound loss, its zero
                           set and FP64 central differences
                           (tests/test_oracle_proposal.py).
* camera rays           -- closed-form directions of chosen pixels, exact
                           axis permutations under 90-degree rotations, unit
                           norms, origins equal to the translation
                           (tests/test_oracle_rays.py).
* remove_small_objects  -- a brute-force 6-neighbour flood fill on random
                           masks, diagonal contact left unconnected
                           (tests/test_occupancy_cpu.py).
* marching tetrahedra   -- a sphere's signed distance: vertices on the sphere,
  (oracle/mesh.py)         a closed, consistently oriented mesh, area and
                           enclosed volume against the closed forms
                           (tests/test_mesh.py).

tests/golden/ holds fixtures written from these functions
(tests/golden/make_golden.py); tests/test_golden.py recomputes them.
"""

from .field import (AttrDict, N_to_reso, N_to_vm_reso, OracleField, cal_n_samples,
                    dct_dict, grid_mapping, make_cfg, setup_params)
from .rays import all_rays, gather_batch, get_ray_directions, get_ray_directions_blender, get_rays
from .render import (OracleRenderer, alpha_grid_sample, composite, eval_sh, mlp_render_fea,
                     positional_encoding, raw2alpha, sample_pdf, sample_point)

__all__ = [
    "AttrDict", "N_to_reso", "N_to_vm_reso", "OracleField", "cal_n_samples",
    "dct_dict", "grid_mapping", "make_cfg", "setup_params", "OracleRenderer",
    "alpha_grid_sample", "composite", "eval_sh", "mlp_render_fea", "positional_encoding",
    "raw2alpha", "sample_pdf", "sample_point", "all_rays", "gather_batch", "get_ray_directions",
    "get_ray_directions_blender", "get_rays",
]
