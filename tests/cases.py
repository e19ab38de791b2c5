"""Matched problem builders for the GPU product and the CPU oracle.

Every case is built twice from the same description: once with
``jaxiga_amd`` (GPU) and once with ``oracle`` (CPU FP64). Point callables are
written once against an array module so they run on torch GPU tensors and on
NumPy arrays alike.
"""

from __future__ import annotations

import dataclasses

import numpy as np
import torch

import jaxiga_amd as jx
from oracle import forms as OF
from oracle import geometry as OG
from oracle import space as OS

UNIT = [[0, 0], [1, 0], [1, 1], [0, 1]]
TWO_PI = 2 * np.pi


def xp(a):
    return torch if isinstance(a, torch.Tensor) else np


def sin_source(x, p=None):
    m = xp(x)
    return 2 * TWO_PI**2 * m.sin(TWO_PI * x[0]) * m.sin(TWO_PI * x[1])


def sin_exact(x):
    m = xp(x)
    return m.sin(TWO_PI * x[0]) * m.sin(TWO_PI * x[1])


def everywhere(x):
    return np.ones(x.shape[0], bool)


def unit_traction(x, n, p=None):
    return n


@dataclasses.dataclass
class Case:
    name: str
    kind: str
    space: object  # jaxiga_amd.FunctionSpace
    ospace: object  # oracle FunctionSpace
    problem: object
    odirichlet: list
    oneumann: list
    source: object
    params: dict
    plane: str = "stress"


def patches(geom, deg, ref):
    if geom == "square":
        return (jx.primitives.quadrilateral(UNIT).elevate(deg).refine(ref),
                OG.quadrilateral(UNIT).elevate(deg).refine(ref))
    if geom == "annulus":
        return (jx.primitives.quarter_annulus(1.0, 2.0).elevate(deg).refine(ref),
                OG.quarter_annulus(1.0, 2.0).elevate(deg).refine(ref))
    if geom == "cube":
        return (jx.primitives.cuboid([0, 0, 0], [1, 1, 1]).elevate(deg).refine(ref),
                OG.cuboid([0, 0, 0], [1, 1, 1]).elevate(deg).refine(ref))
    raise ValueError(geom)


def poisson_case(geom="square", deg=3, ref=3, a0=1.0, source=sin_source, g=0.0):
    P, O = patches(geom, deg, ref)
    V = jx.FunctionSpace(P)
    VO = OS.function_space(O)
    prob = jx.Poisson(V, dirichlet=[jx.DirichletBC(g, where=everywhere)], source=source)
    return Case(f"poisson-{geom}-p{deg}-r{ref}", "poisson", V, VO, prob,
                [OF.DirichletBC(g, everywhere)], [], source, {"a0": a0})


def elasticity_case(deg=3, ref=2, plane="strain"):
    """Quarter annulus, clamped inner edge, unit normal traction on the outer edge."""
    P, O = patches("annulus", deg, ref)
    V = jx.FunctionSpace(P, vec=2)
    VO = OS.function_space(O, vec=2)
    prob = jx.LinearElasticity(V, plane=plane, dirichlet=[jx.DirichletBC(0.0, where="inner")],
                               neumann=[jx.Neumann(unit_traction, where="outer")])
    return Case(f"elasticity-annulus-p{deg}-r{ref}", "elasticity", V, VO, prob,
                [OF.DirichletBC(0.0, "inner")], [OF.Neumann(unit_traction, "outer")], None,
                {"E": 1.0, "nu": 0.3}, plane)
