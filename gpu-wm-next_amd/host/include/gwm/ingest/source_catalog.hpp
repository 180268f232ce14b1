// Reference include path include/gwm/ingest/source_catalog.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_ingest.hpp"
