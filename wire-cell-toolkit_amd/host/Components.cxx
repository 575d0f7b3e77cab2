// Registers the small configurable components (Components.h).
#include "Components.h"
#include "Factory.h"

WIRECELL_HIP_FACTORY(HfFilter, WireCell::Hip::HfFilter)
WIRECELL_HIP_FACTORY(LfFilter, WireCell::Hip::LfFilter)
WIRECELL_HIP_FACTORY(ColdElecResponse, WireCell::Hip::ColdElecResponse)
WIRECELL_HIP_FACTORY(FieldResponse, WireCell::Hip::FieldResponseFile)
WIRECELL_HIP_FACTORY(PerChannelResponse, WireCell::Hip::PerChannelResponse)
WIRECELL_HIP_FACTORY(AnodeChannels, WireCell::Hip::AnodeChannels)
