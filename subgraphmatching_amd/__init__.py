"""MI355X (gfx950) engine for the filter -> order -> enumerate path of subgraph matching
(labelled or not), with the call surface of Orrinn/SubgraphMatching (inc/api.h, filter.h,
order.h, explore.h).  See ``subgraphmatching_amd.api`` for the mirrored entry points."""
from .graph import Graph, Query, load_cached, load_csr, load_text, save_csr, save_text  # noqa: F401
from . import generators  # noqa: F401
from .api import (Batch, CandidateParam, Config, Context, ExploreResult, Filter_CECI, Filter_CFL,  # noqa: F401
                  Filter_DAF, Filter_GQL, Filter_NLF, Order_CECI, Order_CFL, Order_GQL, Order_GuP, Order_QSI,
                  Order_RI, Profiler, query_kcore, ReuseParam, default_context, enumerate_batch, enumerate_query,
                  explore_Intersection_All, explore_Intersection_Cache_DAF_Filter, filter_batch,
                  plan_export)

__version__ = "0.1.0"
