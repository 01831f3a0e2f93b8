"""Atom data for the level-fraction head: per species the level count, the particle mass the total is referred to,
and MULTI3D's atom data (statistical weights g, level energies ev [eV], ionisation stages ion).

The values are data the checkpoints were trained with, so they are the reference's (config.py:150-220); nothing
else of config.py is on the prediction path.
"""
import math

ATOMS = {
    "H": dict(
        levels=6,
        log10_particle_mass=math.log10(2.2697e-27),       # kg (config.py:181)
        multi3d_lte=dict(g=[2.0, 8.0, 18.0, 32.0, 50.0, 1.0], ion=[1, 1, 1, 1, 1, 2],
                         ev=[0.0, 10.198673740091325, 12.087318311066094, 12.748311674749516, 13.054262568222963,
                             13.60794573029334]),
    ),
    "CA": dict(
        levels=6,
        log10_particle_mass=math.log10(1.0375e-21),       # kg (config.py:214)
        multi3d_lte=dict(g=[2.0, 4.0, 6.0, 2.0, 4.0, 1.0], ion=[2, 2, 2, 2, 2, 3],
                         ev=[0.0, 1.69242903005873, 1.6999489817497997, 3.12337806212316, 3.151018087686871,
                             11.87181686885173]),
    ),
}

# The final model's constructor defaults (config.py:294-311); dropout is inactive in eval mode and omitted.
MODEL_CONFIG = dict(in_channels=6, out_channels=6, width=96, n_layers=4, hidden_spectral=128, hidden_vertical=128,
                    checkpoint_blocks=True, freq_mix_rank=8, z_reference_length=40.0, z_min_wavelength=0.04,
                    z_feature_wavelength=0.3)


def species(names):
    """The species list FFNO3D takes, for atom names such as ["H", "CA"]."""
    return [dict(levels=ATOMS[a]["levels"], log10_particle_mass=ATOMS[a]["log10_particle_mass"],
                 multi3d_lte=ATOMS[a]["multi3d_lte"]) for a in names]
