"""The reference's NLTE response case (tests/golden/nlterin.npz, the Ca II
atom of the cfg0 column) in the layouts of snapi_nlte_response_system."""
from snapi_amd import workloads


def load():
    return workloads.nlte_system_case()
