// Reference include path include/gwm/ingest/analysis_state.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_ingest.hpp"
