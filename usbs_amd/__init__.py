"""MI355X-native (gfx950) backend for the USBS hot path."""
