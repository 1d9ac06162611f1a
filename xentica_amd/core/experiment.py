"""
Experiments: class-level settings of a model run (``xentica/core/experiment.py``).

Every non-callable attribute of an experiment class (``word``, ``size``,
``seed``, ``zoom``, ``pos``, parameter values, colour settings ...) is
copied onto the model at construction (``xentica/core/base.py:641-649``).

Example::

    class MyExperiment(core.Experiment):
        word = "My Special String"
        size = (640, 360, )
        zoom = 3
        pos = [0, 0]
        seed = seeds.patterns.BigBang(
            pos=(320, 180), size=(100, 100),
            vals={"state": seeds.random.RandInt(0, 1)})
"""

__all__ = ['Experiment', ]


class Experiment:
    """The base class for all experiments."""
