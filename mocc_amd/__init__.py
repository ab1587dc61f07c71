"""MI355X-native 2-D Method-of-Characteristics sweeper, source update and CMFD
for mocc input decks (see include/mocgpu.h for the C ABI)."""
