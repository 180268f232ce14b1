// Reference include path include/gwm/ingest/runtime_case.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_ingest.hpp"
