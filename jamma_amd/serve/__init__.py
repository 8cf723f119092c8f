"""Continuous-batching association serving over HBM-resident cohorts."""

from jamma_amd.serve.server import Request, ResidentCohort, Server

__all__ = ["Request", "ResidentCohort", "Server"]
