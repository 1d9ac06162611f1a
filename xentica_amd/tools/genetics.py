"""
Genome crossover for evolutionary models (``xentica/tools/genetics.py``).

Builds the DSL statements of a stochastic gene-by-gene crossover: a random
start gene, then for every gene a uniformly chosen donor among the live
genomes, an upper limit on active genes and an optional mutation.  Every
``.uniform`` read advances the model's LCG property once, so the number
and order of draws are part of the model's semantics and match
``genetics.py:7-56``.
"""
from xentica_amd.core.variables import IntegerVariable
from xentica_amd.tools import xmath

__all__ = ['genome_crossover', ]


def genome_crossover(state, num_genes, *genomes, max_genes=None,
                     mutation_prob=0, rng_name="rng"):
    """
    Cross given genomes over in a stochastic way.

    :param state: container holding the model's properties (e.g. ``main``)
    :param num_genes: genome length in bits
    :param genomes: integer genomes to cross over
    :param max_genes: upper limit of set genes in the result
    :param mutation_prob: probability of a single gene's mutation
    :param rng_name: name of the ``RandomProperty`` in ``state``
    :returns: the resulting genome (a DSL variable)
    """
    max_genes = max_genes or num_genes
    gene_choose = IntegerVariable()
    new_genome = IntegerVariable()
    num_genomes = IntegerVariable()
    num_active = IntegerVariable()
    new_gene = IntegerVariable()
    rand_val = getattr(state, rng_name).uniform
    start_gene = IntegerVariable()
    start_gene *= 0
    start_gene += xmath.int(rand_val * num_genes) % len(genomes)
    for gene in range(num_genes):
        gene_choose *= 0
        num_genomes *= 0
        gene = (gene + start_gene) % num_genes
        for genome in genomes:
            gene_choose += ((genome >> gene) & 1 & (genome > 0)) << num_genomes
            num_genomes += (genome > 0)
        rand_val = getattr(state, rng_name).uniform
        winner_gene = xmath.int(rand_val * num_genomes)
        new_gene *= 0
        new_gene += ((gene_choose >> winner_gene) & 1)
        num_active += new_gene
        new_gene *= num_active <= max_genes
        is_mutated = 0
        if mutation_prob > 0:
            is_mutated = getattr(state, rng_name).uniform < mutation_prob
            is_mutated = is_mutated * (gene_choose > 0)
        new_genome += (new_gene ^ is_mutated) << gene
    return new_genome
