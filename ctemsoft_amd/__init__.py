"""MI355X-native EBSD master-pattern Bloch-wave path (CTEMsoft CTEMEBSDmaster beam loop)."""
