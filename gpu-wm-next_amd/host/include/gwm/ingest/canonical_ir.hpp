// Reference include path include/gwm/ingest/canonical_ir.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_ingest.hpp"
