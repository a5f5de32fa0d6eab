"""MI355X (gfx950) engine for the filter -> order -> enumerate path of subgraph matching
(labelled or not), with the call surface of Orrinn/SubgraphMatching (inc/api.h, filter.h,
order.h, explore.h).  See ``subgraphmatching_amd.api`` for the mirrored entry points."""
from .graph import Graph, Query, load_cached, load_csr, load_text, save_csr, save_text  # noqa: F401
from . import generators  # noqa: F401
from .api import (Batch, CandidateParam, Config, Context, DAFStartVertex, ExploreResult, Filter_CECI,  # noqa: F401
                  Filter_CFL, Filter_DAF, Filter_GQL, Filter_NLF, Order_CECI, Order_CFL, Order_GQL, Order_GuP, Order_QSI,
                  Order_RI, Profiler, query_kcore, ReuseParam, default_context, enumerate_batch, enumerate_query,
                  explore_Intersection_All, explore_Intersection_Cache_DAF_Filter,
                  explore_Intersection_DynamicOrder_DAF, explore_Intersection_DynamicOrder_VEQ, computeWeightArray,
                  filter_batch, gup_guards, plan_export)

__version__ = "0.1.0"
