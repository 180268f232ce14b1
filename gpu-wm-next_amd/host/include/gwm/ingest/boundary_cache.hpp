// Reference include path include/gwm/ingest/boundary_cache.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_ingest.hpp"
