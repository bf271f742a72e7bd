"""MI355X-native SumCheck prover: Python host layer over libsumcheck_hip.so.

Mirrors the reference package ``zkduel`` (src/zkduel/__init__.py): ``sumcheck``,
``available_backends``, ``get_workload`` and ``WORKLOADS``, plus the device objects
(``Context``, ``Proof``, ``Composition``) and the field-sweep driver (``driver.time_one``).
"""

from .api import available_backends, sumcheck
from .fields import FIELD_SPECS, FieldSpec, get_field
from .prover import Composition, Context, Proof, default_context
from .workloads import SPECS, WORKLOADS, get_spec, get_workload

__all__ = [
    "FIELD_SPECS",
    "FieldSpec",
    "SPECS",
    "WORKLOADS",
    "Composition",
    "Context",
    "Proof",
    "available_backends",
    "default_context",
    "get_field",
    "get_spec",
    "get_workload",
    "sumcheck",
]
